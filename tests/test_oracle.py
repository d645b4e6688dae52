"""Pins of the CPU oracle (oracle/spect_oracle.cpp), independent of its own code.

Each check fixes a part of the restatement by something other than re-typing
its formula: the reference's own test families (test/*.jl), closed forms, and
independent library routines (numpy rot90, scipy.ndimage).  The reference
ships no golden vectors; tests/golden holds FP64-oracle fixtures that pin the
oracle against regressions and serve as the GPU references.
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.ndimage as ndi

from conftest import rand_sym_psfs

HERE = os.path.dirname(os.path.abspath(__file__))


def f32(x):
    return float(np.float32(x))


# ---------------------------------------------------------------------------
# psf_gauss known answers (test/psf-gauss.jl:7-31)
# ---------------------------------------------------------------------------
def test_psf_gauss_known_answers(O):
    ny, px, pz = 4, 5, 3
    psf = O.psf_gauss(ny, px, pz, fwhm=np.zeros(ny))
    tmp = np.zeros((px, pz))
    tmp[(px + 1) // 2 - 1, (pz + 1) // 2 - 1] = 1
    assert np.array_equal(psf, np.repeat(tmp[:, :, None], ny, 2))
    psf = O.psf_gauss(ny, px, pz, fwhm_x=np.full(ny, np.inf), fwhm_z=np.zeros(ny))
    tmp = np.zeros((px, pz))
    tmp[:, (pz + 1) // 2 - 1] = 1 / px
    assert np.allclose(psf, np.repeat(tmp[:, :, None], ny, 2))
    psf = O.psf_gauss(128, 11)  # defaults: sums to 1 per plane, symmetric
    assert np.allclose(psf.sum(axis=(0, 1)), 1, atol=1e-6)
    assert np.array_equal(psf, psf[::-1, ::-1])


def test_psf_gauss_fwhm(O):
    """FWHM semantics: the continuous Gaussian of FWHM f is 1/2 at r = f/2."""
    psf = O.psf_gauss(1, 41, 1, fwhm=[6.0], dtype=np.float64)[:, 0, 0]
    sigma = 6.0 / math.sqrt(math.log(256))
    assert math.isclose(math.exp(-0.5 * (3.0 / sigma) ** 2), 0.5, rel_tol=1e-12)
    assert np.isclose(psf[20 + 3] / psf[20], math.exp(-0.5 * (3 / sigma) ** 2))


# ---------------------------------------------------------------------------
# quarter turns (test/rotatez.jl:7-46) against numpy
# ---------------------------------------------------------------------------
def test_rot_f90_vs_numpy(O):
    rng = np.random.default_rng(0)
    A = np.asfortranarray(rng.random((20, 20), dtype=np.float32))
    assert np.array_equal(O.rot_f90(A, 1), np.rot90(A, 1))   # rotl90
    assert np.array_equal(O.rot_f90(A, 3), np.rot90(A, -1))  # rotr90
    assert np.array_equal(O.rot_f90(A, 2), np.rot90(A, 2))   # rot180
    assert np.array_equal(O.rot_f90(A, 0), A)
    assert np.array_equal(O.rot_f90(A, 1, adj=True), np.rot90(A, -1))
    assert np.array_equal(O.rot_f90(A, 3, adj=True), np.rot90(A, 1))
    for m in (-1, 4):
        with pytest.raises(ValueError, match="invalid m!"):
            O.rot_f90(A, m)


# ---------------------------------------------------------------------------
# linearinterp! tables (src/rotatez.jl:21-51) = linear interpolation
# ---------------------------------------------------------------------------
def _apply_table(Cv, J, x):
    n = len(Cv) // 2
    return np.array([Cv[2 * k] * x[J[2 * k] - 1] + Cv[2 * k + 1] * x[J[2 * k + 1] - 1] for k in range(n)])


@pytest.mark.parametrize("shift", [0.37, -0.81, 2.5, -3.25, 11.9, -7.02])
def test_linearinterp_is_linear_interpolation(O, shift):
    n = 40
    x = np.random.default_rng(1).standard_normal(n)
    s, e = 1 + shift, n + shift
    Cv, J = O.linearinterp(s, e, n)
    got = _apply_table(Cv.astype(np.float64), J, x)
    # np.interp clamps outside [1, n] to the edge values, like the index clamps
    want = np.interp(np.arange(1, n + 1) + shift, np.arange(1, n + 1), x)
    assert np.allclose(got, want, rtol=0, atol=1e-6 * np.abs(x).max())


def test_linearinterp_integer_start(O):
    """Integer s: exact shift when e <= len, and the reference's +1 shift
    (taps (0, 1) at floor(s)+k) when e > len."""
    n = 12
    x = np.arange(1.0, n + 1)
    Cv, J = O.linearinterp(-2.0, n - 3.0, n)  # e <= n
    assert np.array_equal(_apply_table(Cv, J, x), np.clip(np.arange(1, n + 1) - 3, 1, n))
    Cv, J = O.linearinterp(3.0, n + 2.0, n)  # e > n
    assert np.array_equal(_apply_table(Cv, J, x), np.clip(np.arange(1, n + 1) + 3, 1, n))


# ---------------------------------------------------------------------------
# rotation
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("method", ["one", "two"])
def test_rotate_identity_and_quarter_turns(O, method):
    rng = np.random.default_rng(2)
    img = np.asfortranarray(rng.random((20, 20), dtype=np.float32))
    assert np.array_equal(O.imrotate(img, 0.0, method), img)
    for k in (1, 2, 3):
        r = O.imrotate(img, f32(k * math.pi / 2), method)
        if method == "one":  # θ ≈ mπ/2 takes the exact quadrant path (src/rotatez.jl:258-261)
            assert np.array_equal(r, np.rot90(img, k))
        else:
            assert np.abs(r - np.rot90(img, k)).max() <= 1e-5


@pytest.mark.parametrize("method", ["one", "two"])
def test_rotate_dense_adjoint(O, method):
    """Matrix(A') ≈ Matrix(A)' at 16x16 for θ = (0:6)/7·2π (test/adjoint-rotate.jl:43-58)."""
    M = 16
    eye = np.eye(M * M).reshape(M, M, M * M, order="F")
    for k in range(7):
        th = f32(k / 7 * 2 * math.pi)
        A = O.imrotate(eye, th, method, dtype=np.float64).reshape(M * M, M * M, order="F")
        At = O.imrotate(eye, th, method, adj=True, dtype=np.float64).reshape(M * M, M * M, order="F")
        assert np.allclose(A.T, At, atol=1e-14)


@pytest.mark.parametrize("method", ["one", "two"])
def test_rotate_bounds(O, method):
    """test/adjoint-rotate.jl:28-37."""
    rng = np.random.default_rng(3)
    img = np.asfortranarray(rng.random((20, 20, 4), dtype=np.float32))
    assert O.imrotate(img, math.pi / 6, method).max() <= 1.01
    assert O.imrotate(img, math.pi / 6, method, adj=True).max() <= 1.5


def test_rotation_direction_pin(O):
    """:two's map is pinned by agreement with :one (fully specified by the
    reference source); the opposite rotation sense disagrees grossly."""
    N = 32
    x, y = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    img = np.exp(-((x - 12) ** 2 / 30 + (y - 18) ** 2 / 10)).astype(np.float32)
    for th in (0.3, 1.2, 2.5, 4.0):
        a = O.imrotate(img, th, "one")
        assert np.linalg.norm(a - O.imrotate(img, th, "two")) / np.linalg.norm(a) < 0.03
        assert np.linalg.norm(a - O.imrotate(img, 2 * math.pi - th, "two")) / np.linalg.norm(a) > 0.5


def test_bilinear_vs_scipy(O):
    """:two = bilinear sampling of the zero-padded slice at c + Rot(2π-θ)(i - c)
    (src/rotatez.jl:348-357), checked against scipy.ndimage.map_coordinates."""
    N = 24
    pad = O.padsize(N)
    Np = N + 2 * pad
    c = (Np + 1) / 2
    rng = np.random.default_rng(4)
    img = rng.random((N, N))
    padded = np.zeros((Np, Np))
    padded[pad:pad + N, pad:pad + N] = img
    for th in (0.4, 2.2, 3.9, 5.5):
        phi = 2 * math.pi - th
        i1, i2 = np.meshgrid(np.arange(pad + 1, pad + N + 1), np.arange(pad + 1, pad + N + 1), indexing="ij")
        x1 = c + math.cos(phi) * (i1 - c) - math.sin(phi) * (i2 - c)
        x2 = c + math.sin(phi) * (i1 - c) + math.cos(phi) * (i2 - c)
        want = ndi.map_coordinates(padded, [x1 - 1, x2 - 1], order=1, mode="constant", cval=0.0)
        got = O.imrotate(img, th, "two", dtype=np.float64)
        assert np.abs(got - want).max() < 1e-10


# ---------------------------------------------------------------------------
# PSF convolution (src/fft_convolve.jl)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nx,nz,px,pz", [(20, 14, 5, 5), (12, 8, 7, 7), (33, 17, 9, 3), (16, 1, 5, 1), (9, 9, 1, 1)])
def test_fft_conv_closed_form(O, nx, nz, px, pz):
    """Literal FFT restatement = direct clamped convolution = scipy's
    convolve(mode='nearest') (replicate padding)."""
    rng = np.random.default_rng(nx + 7 * pz)
    img = rng.standard_normal((nx, nz))
    ker = rng.random((px, pz))
    ker = ker + ker[::-1, ::-1]
    ker /= ker.sum()
    f = O.fft_conv(img, ker, False, "fft", np.float64)
    d = O.fft_conv(img, ker, False, "direct", np.float64)
    s = ndi.convolve(img, ker, mode="nearest")
    assert np.abs(f - d).max() < 1e-12
    assert np.abs(d - s).max() < 1e-12
    fa = O.fft_conv(img, ker, True, "fft", np.float64)
    da = O.fft_conv(img, ker, True, "direct", np.float64)
    assert np.abs(fa - da).max() < 1e-12


def test_fft_conv_dense_adjoint(O):
    """test/fftconv.jl:51-68: Matrix(A') ≈ Matrix(A)' at 20x14 with 5x5 kernels."""
    M, N = 20, 14
    rng = np.random.default_rng(5)
    for _ in range(2):
        ker = rng.random((5, 5))
        ker = ker + ker[::-1, ::-1]
        ker /= ker.sum()
        eye = np.eye(M * N)
        A = np.stack([O.fft_conv(e.reshape(M, N, order="F"), ker, False, "fft", np.float64).ravel(order="F")
                      for e in eye], 1)
        At = np.stack([O.fft_conv(e.reshape(M, N, order="F"), ker, True, "fft", np.float64).ravel(order="F")
                       for e in eye], 1)
        assert np.allclose(A.T, At, atol=1e-14)


def test_fft_conv_bounds(O):
    """test/fftconv.jl:32-48: box kernel of rand <= 1, adjoint <= 1.5."""
    rng = np.random.default_rng(6)
    img = rng.random((12, 10)).astype(np.float32)
    ker = np.ones((5, 3), np.float32) / 15
    assert O.fft_conv(img, ker, False).max() <= 1
    assert O.fft_conv(img, ker, True).max() <= 1.5


def test_fft_conv_f32_vs_f64(O):
    rng = np.random.default_rng(7)
    img = rng.standard_normal((128, 80))
    ker = np.outer(*(2 * [np.exp(-0.5 * (np.arange(-18, 19) / 2.0) ** 2)]))
    ker /= ker.sum()
    r = O.fft_conv(img, ker, False, "direct", np.float64)
    g = O.fft_conv(img, ker, False, "fft", np.float32)
    assert np.abs(g - r).max() / np.abs(r).max() < 2e-6


# ---------------------------------------------------------------------------
# attenuation (src/project.jl:103-113)
# ---------------------------------------------------------------------------
def test_attenuation_closed_form(O):
    nx, nz, mu, dy = 16, 4, 0.15, 0.48
    psfs = np.ones((1, 1, nx, 4), np.float32)
    ang = np.zeros(4)
    P = O.Problem(np.zeros((nx, nx, nz), np.float32), psfs, dy, ang, "two")
    assert np.array_equal(P.atten(0), np.ones((nx, nx, nz), np.float32))
    P = O.Problem(np.full((nx, nx, nz), mu), psfs, dy, ang, "two", np.float64)
    a = P.atten(0)  # θ = 0: no rotation; plane y (1-based) sees exp(-mu dy (y - 1/2))
    want = np.exp(-mu * dy * (np.arange(1, nx + 1) - 0.5))
    assert np.allclose(a, want[None, :, None], rtol=1e-14)


# ---------------------------------------------------------------------------
# projector
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("method", ["one", "two"])
def test_project_line_integrals(O, method):
    """mu = 0, PSF = δ: view 0 (θ = 0) is Σ_y image and view at θ = π the
    flipped sum (exact quarter-turn path for :one)."""
    nx, nz, nview = 16, 5, 4
    rng = np.random.default_rng(8)
    img = rng.random((nx, nx, nz))
    psfs = np.ones((1, 1, nx, nview))
    # Float64 angles: the FP64 oracle's quadrant test uses rtol sqrt(eps(Float64))
    ang = np.array([0.0, math.pi / 2, math.pi, 3 * math.pi / 2])
    P = O.Problem(np.zeros((nx, nx, nz)), psfs, 1.0, ang, method, np.float64, "direct")
    v = P.project(img)
    assert np.allclose(v[:, :, 0], img.sum(axis=1), rtol=1e-12)
    tol = 1e-12 if method == "one" else 1e-5
    assert np.allclose(v[:, :, 2], img.sum(axis=1)[::-1, :], rtol=tol, atol=tol)


@pytest.mark.parametrize("method", ["one", "two"])
def test_project_dot_adjoint(O, method):
    """⟨y, Ax⟩ = ⟨A'y, x⟩ in Float64 at 32x32x20, 30 views, 5x3 PSF (test/adjoint-project.jl:52-107)."""
    nx, nz, nview = 32, 20, 30
    rng = np.random.default_rng(9)
    mumap = rng.random((nx, nx, nz))
    psfs = rand_sym_psfs(rng, 5, 3, nx, nview).astype(np.float64)
    ang = np.arange(nview) / nview * f32(2 * math.pi)
    P = O.Problem(mumap, psfs, 4.7952, ang, method, np.float64, "fft")
    x = rng.standard_normal((nx, nx, nz))
    y = rng.standard_normal((nx, nz, nview))
    a, b = np.vdot(y, P.project(x)), np.vdot(x, P.backproject(y))
    assert abs(a - b) <= 1e-10 * abs(a)


@pytest.mark.parametrize("method", ["one", "two"])
def test_project_dense_adjoint(O, method):
    """Matrix(A)' ≈ Matrix(A') at 8x8x6, 7 views, 5x3 PSF (test/adjoint-project.jl:12-49)."""
    nx, nz, nview = 8, 6, 7
    rng = np.random.default_rng(10)
    mumap = rng.random((nx, nx, nz))
    psfs = rand_sym_psfs(rng, 5, 3, nx, nview).astype(np.float64)
    ang = np.arange(nview) / nview * f32(2 * math.pi)
    P = O.Problem(mumap, psfs, 4.7952, ang, method, np.float64, "direct")
    N, M = nx * nx * nz, nx * nz * nview
    A = np.stack([P.project(e.reshape(nx, nx, nz, order="F")).ravel(order="F") for e in np.eye(N)], 1)
    At = np.stack([P.backproject(e.reshape(nx, nz, nview, order="F")).ravel(order="F") for e in np.eye(M)], 1)
    assert np.allclose(A.T, At, atol=1e-13)


def test_project_fp32_vs_fp64(O):
    nx, nz, nview = 32, 16, 12
    rng = np.random.default_rng(11)
    mumap = (0.1 * rng.random((nx, nx, nz))).astype(np.float32)
    psfs = np.repeat(O.psf_gauss(nx, 7, 7, 1.0, 3.0)[:, :, :, None], nview, 3)
    img = rng.random((nx, nx, nz)).astype(np.float32)
    ang = np.arange(nview) / nview * f32(2 * math.pi)
    r32 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float32, "fft").project(img)
    r64 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float64, "direct").project(img.astype(np.float64))
    for k in range(nview):
        assert np.abs(r32[:, :, k] - r64[:, :, k]).max() / np.abs(r64[:, :, k]).max() < 2e-5


# ---------------------------------------------------------------------------
# ML-EM / OS-EM (src/ml-os-em.jl)
# ---------------------------------------------------------------------------
def _em_problem(O):
    nx, nz, nview = 16, 6, 8
    rng = np.random.default_rng(12)
    mumap = 0.02 * rng.random((nx, nx, nz))
    psfs = np.repeat(O.psf_gauss(nx, 5, 3, 1.0, 2.0, dtype=np.float64)[:, :, :, None], nview, 3)
    ang = np.arange(nview) / nview * f32(2 * math.pi)
    P = O.Problem(mumap, psfs, 0.48, ang, "two", np.float64, "direct")
    x = rng.random((nx, nx, nz)) + 0.5
    ybar = P.project(x)
    b = np.full_like(ybar, 0.1 * ybar.mean())
    y = rng.poisson(50 * (ybar + b)) / 50
    return P, y, b


def test_em_identity(O):
    """Each ML-EM iterate satisfies Σ_j asum_j x_j^{k+1} = Σ_i y_i (Ax^k)_i / ((Ax^k)_i + b_i)."""
    P, y, b = _em_problem(O)
    x0 = np.ones((P.nx, P.ny, P.nz))
    asum = P.backproject(np.ones_like(y))
    x1 = P.mlem(x0, y, b, 1)
    x2 = P.mlem(x0, y, b, 2)
    for xk, xk1 in ((x0, x1), (x1, x2)):
        Ax = P.project(xk)
        assert np.isclose(np.sum(asum * xk1), np.sum(y * Ax / (Ax + b)), rtol=1e-12)


def test_osem_one_block_is_mlem(O):
    P, y, b = _em_problem(O)
    x0 = np.ones((P.nx, P.ny, P.nz))
    assert np.allclose(P.osem(x0, y, b, 1, 3), P.mlem(x0, y, b, 3), rtol=1e-13)


def test_em_monotone_likelihood(O):
    """ML-EM does not decrease the Poisson log-likelihood."""
    P, y, b = _em_problem(O)
    x = np.ones((P.nx, P.ny, P.nz))

    def ll(x):
        yb = P.project(x) + b
        return float(np.sum(y * np.log(yb) - yb))

    last = ll(x)
    for k in range(1, 4):
        cur = ll(P.mlem(x, y, b, k))
        assert cur >= last - 1e-9
        last = cur


# ---------------------------------------------------------------------------
# default angles and golden fixtures
# ---------------------------------------------------------------------------
def test_default_viewangle_exact(S):
    """(0:nview-1)/nview*Float32(2π) element k = fl32(k·fl32(2π)/nview) (src/spectplan.jl:63)."""
    for nview in (7, 60, 128):
        va = S.default_viewangle(nview)
        twopi = Fraction(float(np.float32(2 * math.pi)))
        for k in range(nview):
            exact = twopi * k / nview
            lo, hi = float(np.float32(float(exact))), None
            # round-to-nearest of the rational value
            cand = [np.float32(float(exact))]
            cand += [np.nextafter(cand[0], np.float32(10)), np.nextafter(cand[0], np.float32(-10))]
            best = min(cand, key=lambda c: abs(Fraction(float(c)) - exact))
            assert va[k] == best, (nview, k)


GOLDEN = os.path.join(HERE, "golden", "spect_small.npz")


def test_oracle_matches_golden(O):
    """FP64-oracle fixtures written by tests/golden/make_golden.py."""
    import sys

    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_golden as G

    ref = np.load(GOLDEN, allow_pickle=False)
    got = G.compute(O)
    for k in ref.files:
        assert np.allclose(got[k], ref[k], rtol=1e-12, atol=1e-12 * np.abs(ref[k]).max()), k


@pytest.mark.parametrize("method", ["one", "two"])
def test_gradient_finite_differences(O, method):
    """The oracle's gradients of L = ||A(x, mu) - y||^2 against central finite
    differences of its own FP64 loss: w.r.t. x (L is quadratic in x) and
    w.r.t. the mu-map (through the rotated exponential cumulative sum)."""
    nx, nz, nview, dy = 8, 4, 3, 0.7
    rng = np.random.default_rng(61)
    mumap = np.asfortranarray(0.3 * rng.random((nx, nx, nz)))
    psfs = np.asfortranarray(rand_sym_psfs(rng, 3, 3, nx, nview).astype(np.float64))
    ang = np.array([0.0, 0.7, 2.9])
    x = np.asfortranarray(rng.random((nx, nx, nz)))
    y = np.asfortranarray(rng.random((nx, nz, nview)))
    P = O.Problem(mumap, psfs, dy, ang, method, np.float64, "direct")
    loss, gx, gmu = P.grad(x, y)
    r = P.project(x) - y
    assert abs(loss - float(np.sum(r * r))) <= 1e-12 * loss
    h = 1e-6
    for k in [(0, 0, 0), (3, 4, 1), (7, 2, 3), (5, 5, 2)]:
        e = np.zeros_like(x)
        e[k] = h
        fd = (P.grad(x + e, y)[0] - P.grad(x - e, y)[0]) / (2 * h)
        assert abs(fd - gx[k]) <= 1e-6 * max(1.0, abs(gx[k])), (k, fd, gx[k])
        mp = np.asfortranarray(mumap.copy())
        mm = np.asfortranarray(mumap.copy())
        mp[k] += h
        mm[k] -= h
        lp = O.Problem(mp, psfs, dy, ang, method, np.float64, "direct").grad(x, y)[0]
        lm = O.Problem(mm, psfs, dy, ang, method, np.float64, "direct").grad(x, y)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gmu[k]) <= 1e-6 * max(1.0, abs(gmu[k])), (k, fd, gmu[k])
    # a subset of views: the same as the full gradient of the restricted loss
    idx = np.array([2, 0], np.int32)
    ls, gxs, _ = P.grad(x, np.asfortranarray(y[:, :, idx]), idx)
    r = P.project(x, idx) - y[:, :, idx]
    assert abs(ls - float(np.sum(r * r))) <= 1e-12 * ls
    assert np.allclose(gxs, P.backproject(2 * r, idx), rtol=1e-12, atol=1e-14)


def test_osem_uneven_blocks_composition(O):
    """OS-EM over strided blocks nb:nblocks:nview when nblocks does not divide
    nview (the batched use): equal to the update composed from the pinned
    project / backproject, block by block."""
    nx, nz, nview, nblocks, niter = 8, 4, 5, 2, 2
    rng = np.random.default_rng(67)
    mumap = np.asfortranarray(0.1 * rng.random((nx, nx, nz)))
    psfs = np.asfortranarray(rand_sym_psfs(rng, 3, 3, nx, nview).astype(np.float64))
    P = O.Problem(mumap, psfs, 0.5, np.arange(nview) * 2 * np.pi / nview, "two", np.float64, "direct")
    y = np.asfortranarray(1 + rng.random((nx, nz, nview)))
    b = np.asfortranarray(0.1 + np.zeros((nx, nz, nview)))
    x0 = np.asfortranarray(np.ones((nx, nx, nz)))
    blocks = [np.arange(k, nview, nblocks, dtype=np.int32) for k in range(nblocks)]
    asum = []
    for idx in blocks:
        a = P.backproject(np.asfortranarray(np.ones((nx, nz, len(idx)))), idx)
        a[a == 0] = np.inf
        asum.append(a)
    x = x0.copy()
    for _ in range(niter):
        for idx, a in zip(blocks, asum):
            ratio = np.asfortranarray(y[:, :, idx] / (P.project(x, idx) + b[:, :, idx]))
            x = np.asfortranarray(x * (P.backproject(ratio, idx) / a))
    assert np.allclose(P.osem(x0, y, b, nblocks, niter), x, rtol=1e-12, atol=0)
