"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Tolerances (floating point, FP32 on the device):
  * rotations, convolutions: max|g-r|/max|r| <= 2e-6 vs the FP32 oracle
  * projections: per-view max-relative <= 2e-5 vs the FP32 oracle
  * adjoint identity <y, Ax> = <A'y, x>: relative <= 1e-5
  * ML-EM / OS-EM: relative L2 <= 1e-4 vs the FP64 oracle
"""
import math

import numpy as np
import pytest

from conftest import maxrel, per_view_maxrel, rand_sym_psfs

pytestmark = pytest.mark.gpu

ROT_TOL = 2e-6
PROJ_TOL = 2e-5


def angles32(nview):
    twopi = float(np.float32(2 * math.pi))
    return np.array([np.float32(k * twopi / nview) for k in range(nview)], np.float32)


# ---------------------------------------------------------------------------
# rotation
# ---------------------------------------------------------------------------
ANGLES = [0.0, 0.3, float(np.float32(math.pi / 6)), float(np.float32(math.pi / 2)), 2.0,
          float(np.float32(math.pi)), 4.0, float(np.float32(3 * math.pi / 2)), 5.9, float(np.float32(2 * math.pi))]


@pytest.mark.parametrize("method", ["one", "two"])
@pytest.mark.parametrize("n", [16, 20, 64, 128])
def test_imrotate_vs_oracle(gpu, O, method, n):
    rng = np.random.default_rng(n)
    img = np.asfortranarray(rng.random((n, n, 3), dtype=np.float32))
    for th in ANGLES:
        th = float(np.float32(th))
        g = gpu.imrotate(img, th, method=method)
        r = O.imrotate(img, th, method)
        assert maxrel(g, r) <= ROT_TOL, (method, n, th, maxrel(g, r))
        ga = gpu.imrotate_adj(img, th, method=method)
        ra = O.imrotate(img, th, method, adj=True)
        assert maxrel(ga, ra) <= ROT_TOL, ("adj", method, n, th, maxrel(ga, ra))


@pytest.mark.parametrize("n", [139, 140])
def test_imrotate_one_largest(gpu, O, n):
    """:one at the largest padded sizes: Np = 199 keeps its interpolation
    tables in LDS beside the slice, Np = 200 reads them from device memory."""
    rng = np.random.default_rng(n)
    img = np.asfortranarray(rng.random((n, n, 2), dtype=np.float32))
    for th in (0.3, 2.0, 4.0, 5.9):
        th = float(np.float32(th))
        g = gpu.imrotate(img, th, method="one")
        r = O.imrotate(img, th, "one")
        assert maxrel(g, r) <= ROT_TOL, (n, th, maxrel(g, r))
        ga = gpu.imrotate_adj(img, th, method="one")
        ra = O.imrotate(img, th, "one", adj=True)
        assert maxrel(ga, ra) <= ROT_TOL, ("adj", n, th, maxrel(ga, ra))


@pytest.mark.parametrize("method", ["one", "two"])
def test_imrotate_dense_adjoint(gpu, method):
    """Matrix(A') == Matrix(A)' at 16x16, 7 angles (test/adjoint-rotate.jl:43-58):
    all 256 unit images rotated in one 3-D call."""
    M = 16
    eye = np.asfortranarray(np.eye(M * M, dtype=np.float32).reshape(M, M, M * M, order="F"))
    for k in range(7):
        th = float(np.float32(k / 7 * 2 * math.pi))
        A = gpu.imrotate(eye, th, method=method).reshape(M * M, M * M, order="F")
        At = gpu.imrotate_adj(eye, th, method=method).reshape(M * M, M * M, order="F")
        assert np.allclose(A.T, At, rtol=1e-6, atol=1e-7), (method, th)


@pytest.mark.parametrize("method", ["one", "two"])
def test_imrotate_bounds(gpu, method):
    """test/adjoint-rotate.jl:28-37: rotated rand has max <= 1.01, adjoint <= 1.5."""
    rng = np.random.default_rng(7)
    img = np.asfortranarray(rng.random((20, 20, 4), dtype=np.float32))
    th = math.pi / 6
    assert gpu.imrotate(img, th, method=method).max() <= 1.01
    assert gpu.imrotate_adj(img, th, method=method).max() <= 1.5


# ---------------------------------------------------------------------------
# PSF convolution
# ---------------------------------------------------------------------------
CONV_CASES = [(20, 14, 5, 5), (12, 8, 7, 7), (12, 10, 5, 3), (64, 32, 15, 15), (128, 80, 37, 37), (7, 1, 3, 1),
              (33, 17, 9, 1), (16, 16, 1, 1),
              # the reference's fft_conv timing plane (time/fft_convolve.jl:9-17): its separable strips would need
              # more threads than the separable kernels' launch bound, so it takes the direct 2-D kernels
              (200, 64, 5, 5)]


@pytest.mark.parametrize("nx,nz,px,pz", CONV_CASES)
@pytest.mark.parametrize("kind", ["separable", "general"])
def test_fft_conv_vs_oracle(gpu, O, nx, nz, px, pz, kind):
    rng = np.random.default_rng(nx * 100 + nz)
    img = np.asfortranarray(rng.standard_normal((nx, nz)).astype(np.float32))
    if kind == "separable":
        u = rng.random(px) + 0.1
        w = rng.random(pz) + 0.1
        u = u + u[::-1]
        w = w + w[::-1]
        ker = np.outer(u, w)
    else:
        ker = rng.random((px, pz))
        ker = ker + ker[::-1, ::-1]
    ker = np.asfortranarray((ker / ker.sum()).astype(np.float32))
    for adj in (False, True):
        r = O.fft_conv(img.astype(np.float64), ker.astype(np.float64), adj, "direct", np.float64)
        g = gpu.fft_conv_adj(img, ker) if adj else gpu.fft_conv(img, ker)
        assert maxrel(g, r) <= ROT_TOL, (adj, maxrel(g, r))


def test_fft_conv_bounds(gpu):
    """test/fftconv.jl:32-48: box kernel keeps rand <= 1, adjoint <= 1.5."""
    rng = np.random.default_rng(3)
    nx, nz, px, pz = 12, 10, 5, 3
    plan = gpu.plan_psf(nx=nx, nz=nz, px=px, pz=pz)
    image3 = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    ker3 = np.ones((px, pz, nx), np.float32) / (px * pz)
    result = np.zeros_like(image3)
    gpu.fft_conv_(result, image3, ker3, plan)
    assert result.max() <= 1
    gpu.fft_conv_adj_(result, image3, ker3, plan)
    assert result.max() <= 1.5
    gpu.fft_conv_adj2_(result, image3[:, 2, :], ker3, plan)
    assert result.max() <= 1.5


def _ker3(rng, px, pz, n, kind):
    ks = []
    for i in range(n):
        if kind == "separable" or (kind == "mixed" and i % 2 == 0):
            u = rng.random(px) + 0.1
            w = rng.random(pz) + 0.1
            k = np.outer(u + u[::-1], w + w[::-1])
        else:
            k = rng.random((px, pz))
            k = k + k[::-1, ::-1]
        ks.append(k / k.sum())
    return np.asfortranarray(np.stack(ks, axis=2).astype(np.float32))


def test_unit_workspace_reuse(gpu, O):
    """The unit operators keep their tables and buffers between calls: a
    sequence that switches kernels, plane sizes, angles and methods back and
    forth (so cached tables are reused, replaced and grown) matches the
    oracle at every call."""
    rng = np.random.default_rng(21)
    kers = {k: _ker3(rng, *k[:2], 1, k[2])[:, :, 0] for k in [(5, 5, "separable"), (5, 5, "general"),
                                                              (9, 3, "separable")]}
    imgs = {s: np.asfortranarray(rng.standard_normal(s).astype(np.float32)) for s in [(20, 14), (40, 24), (16, 16)]}
    seq = [((20, 14), (5, 5, "separable")), ((20, 14), (5, 5, "general")), ((40, 24), (5, 5, "general")),
           ((20, 14), (5, 5, "separable")), ((16, 16), (9, 3, "separable")), ((20, 14), (5, 5, "separable")),
           ((40, 24), (9, 3, "separable"))]
    for shape, kk in seq:
        img, ker = imgs[shape], kers[kk]
        for adj in (False, True):
            r = O.fft_conv(img.astype(np.float64), ker.astype(np.float64), adj, "direct", np.float64)
            g = gpu.fft_conv_adj(img, ker) if adj else gpu.fft_conv(img, ker)
            assert maxrel(g, r) <= ROT_TOL, (shape, kk, adj)
    vols = {n: np.asfortranarray(rng.standard_normal((n, n, 3)).astype(np.float32)) for n in (24, 40)}
    for n, th, method in [(24, 0.3, "two"), (24, 2.0, "two"), (40, 2.0, "two"), (24, 0.3, "one"), (24, 0.3, "two"),
                          (40, 0.3, "one")]:
        th = float(np.float32(th))
        for adj in (False, True):
            g = (gpu.imrotate_adj if adj else gpu.imrotate)(vols[n], th, method=method)
            r = O.imrotate(vols[n], th, method, adj=adj)
            assert maxrel(g, r) <= ROT_TOL, (n, th, method, adj)
    # device arrays, one plane: output apart from the input, and in place
    from spectrecon_amd import _lib as L
    lib = gpu.lib()
    img, ker = imgs[(20, 14)], kers[(5, 5, "general")]
    dk = gpu.DeviceArray.from_host(ker)
    for adj in (0, 1):
        di, do = gpu.DeviceArray.from_host(img), gpu.DeviceArray((20, 14))
        L.check(lib.sr_fft_conv(do.ptr, di.ptr, 20, 14, dk.ptr, 5, 5, adj, L.SR_MEM_DEVICE))
        L.check(lib.sr_fft_conv(di.ptr, di.ptr, 20, 14, dk.ptr, 5, 5, adj, L.SR_MEM_DEVICE))
        np.testing.assert_array_equal(di.to_host(), do.to_host())
        r = O.fft_conv(img.astype(np.float64), ker.astype(np.float64), bool(adj), "direct", np.float64)
        assert maxrel(do.to_host(), r) <= ROT_TOL, ("device", adj)


@pytest.mark.parametrize("nx,ny,nz,px,pz,kind", [(12, 12, 10, 5, 3, "separable"), (20, 7, 14, 5, 5, "general"),
                                                  (64, 32, 32, 15, 15, "mixed"), (33, 3, 1, 9, 1, "separable"),
                                                  (128, 4, 80, 37, 37, "general")])
def test_fft_conv3_forms_vs_oracle(gpu, O, nx, ny, nz, px, pz, kind):
    """The 3-D forms (src/fft_convolve.jl:160-239) in one launch vs the oracle plane by plane (FP64)."""
    rng = np.random.default_rng(nx + 7 * ny + 31 * nz)
    plan = gpu.plan_psf(nx=nx, nz=nz, px=px, pz=pz)
    image3 = np.asfortranarray(rng.standard_normal((nx, ny, nz)).astype(np.float32))
    ker3 = _ker3(rng, px, pz, ny + 1, kind)  # one plane more than needed: only the first ny are read
    for adj in (False, True):
        out = np.full_like(image3, np.nan)
        (gpu.fft_conv_adj_ if adj else gpu.fft_conv_)(out, image3, ker3, plan)
        for iy in range(ny):
            r = O.fft_conv(image3[:, iy, :].astype(np.float64), ker3[:, :, iy].astype(np.float64), adj, "direct",
                           np.float64)
            assert maxrel(out[:, iy, :], r) <= ROT_TOL, (adj, iy, maxrel(out[:, iy, :], r))
        # in place (output aliases image3) gives the same result
        inplace = image3.copy(order="F")
        (gpu.fft_conv_adj_ if adj else gpu.fft_conv_)(inplace, inplace, ker3, plan)
        np.testing.assert_array_equal(inplace, out)
    # fft_conv_adj2!: planes 1..size(image2, 2) of output from one image
    image2 = np.asfortranarray(rng.standard_normal((nx, nz)).astype(np.float32))
    if nz > ny:
        with pytest.raises(IndexError):
            gpu.fft_conv_adj2_(np.zeros_like(image3), image2, ker3, plan)
        return
    out = np.full_like(image3, 7.0)
    gpu.fft_conv_adj2_(out, image2, ker3, plan)
    for iy in range(ny):
        if iy >= nz:
            assert np.all(out[:, iy, :] == 7.0), iy  # not written, as in the reference
            continue
        r = O.fft_conv(image2.astype(np.float64), ker3[:, :, iy].astype(np.float64), True, "direct", np.float64)
        assert maxrel(out[:, iy, :], r) <= ROT_TOL, ("adj2", iy, maxrel(out[:, iy, :], r))


# ---------------------------------------------------------------------------
# projector / back-projector
# ---------------------------------------------------------------------------
def small_problem(seed, nx=16, nz=8, nview=8, px=5, pz=3, separable=False):
    rng = np.random.default_rng(seed)
    mumap = np.asfortranarray((0.05 * rng.random((nx, nx, nz))).astype(np.float32))
    if separable:
        psfs = np.asfortranarray(np.repeat(
            np.asarray(__import__("oracle").psf_gauss(nx, px, pz, 1.0, 3.0), np.float32)[:, :, :, None], nview, 3))
    else:
        psfs = rand_sym_psfs(rng, px, pz, nx, nview)
    image = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    return mumap, psfs, image


@pytest.mark.parametrize("nx", [144, 210])
@pytest.mark.parametrize("separable", [False, True])
def test_large_slices_vs_oracle(gpu, O, nx, separable):
    """:two projection and back-projection of wide slices (x-fastest: the
    gather adjoint for nx > 128; at 210 the slice no longer fits LDS, so the
    forward rotation and the adjoint read it where it lies) against the FP32
    oracle, and the adjoint identity."""
    nz, nview = 4, 3
    mumap, psfs, image = small_problem(17, nx, nz, nview, 5, 5 if separable else 3, separable)
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    assert plan.info()["layout"] == "x-fastest"
    ang = angles32(nview).astype(np.float64)
    P32 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float32, "fft")
    g = gpu.project(image, plan)
    assert per_view_maxrel(g, P32.project(image)) <= PROJ_TOL
    rng = np.random.default_rng(nx)
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    gb = gpu.backproject(views, plan)
    assert maxrel(gb, P32.backproject(views)) <= PROJ_TOL
    a = float(np.vdot(views.astype(np.float64), g.astype(np.float64)))
    b = float(np.vdot(image.astype(np.float64), gb.astype(np.float64)))
    assert abs(a - b) <= 1e-5 * abs(a)


@pytest.mark.parametrize("nx,nz,nview", [(128, 128, 4), (256, 64, 3), (192, 96, 3)])
def test_large_volumes_vs_oracle(gpu, O, nx, nz, nview):
    """Volumes beyond the z-fastest path's planes (x-fastest: the matrix-core
    PSF kernels on the 128x128 and 192x96 planes, 256-wide slices beyond LDS)
    with a 15x15 Gaussian PSF: projection and back-projection against the
    oracle's FP32 direct convolution (its FP32 FFT restatement is itself up
    to 8e-5 from FP64 at the edges of 128-deep planes) and the FP64 oracle,
    the adjoint identity, and an OS-EM pass that stays finite and
    non-negative."""
    rng = np.random.default_rng(nx + nz)
    mumap = np.asfortranarray((0.02 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = np.asfortranarray(np.repeat(np.asarray(O.psf_gauss(nx, 15, 15, 1.0, 4.0), np.float32)[:, :, :, None],
                                       nview, 3))
    print(f"large: {nx}x{nx}x{nz}, {nview} views", flush=True)
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    assert plan.info()["layout"] == "x-fastest"
    image = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    ang = angles32(nview).astype(np.float64)
    P32 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float32, "direct", nthread=16)
    P64 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float64, "direct", nthread=16)
    # (against FP64 the bar is 1e-4: the :two sample coordinates are float32,
    # about 3e-5 of a pixel apart near 256, so the bilinear weights of wide
    # slices carry that error in the reference's arithmetic as in this one;
    # the oracle's FP32 back-projection is itself 3.3e-5 from FP64 at
    # 256x256x64 and 2.1e-5 at 192x192x96, with mu = 0 as well)
    g = gpu.project(image, plan)
    assert per_view_maxrel(g, P32.project(image)) <= PROJ_TOL
    e64 = per_view_maxrel(g, P64.project(image.astype(np.float64)))
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    gb = gpu.backproject(views, plan)
    assert maxrel(gb, P32.backproject(views)) <= PROJ_TOL
    e64b = maxrel(gb, P64.backproject(views.astype(np.float64)))
    print(f"large: vs FP64 project {e64:.2e}, back-project {e64b:.2e}", flush=True)
    assert e64 <= 1e-4 and e64b <= 1e-4
    a = float(np.vdot(views.astype(np.float64), g.astype(np.float64)))
    b = float(np.vdot(image.astype(np.float64), gb.astype(np.float64)))
    assert abs(a - b) <= 1e-5 * abs(a)
    y = g + 0.1
    xh = gpu.osem(np.ones_like(image), y, np.full_like(y, 0.1), gpu.Ablock(plan, 1), niter=1)
    assert np.all(np.isfinite(xh)) and xh.min() >= 0


@pytest.mark.parametrize("method", ["one", "two"])
@pytest.mark.parametrize("mode", ["fast", "mem"])
@pytest.mark.parametrize("separable", [False, True])
def test_project_vs_oracle(gpu, O, method, mode, separable):
    nx, nz, nview = 16, 8, 8
    mumap, psfs, image = small_problem(11, nx, nz, nview, 5, 3 if not separable else 5, separable)
    dy = 0.48
    plan = gpu.SPECTplan(mumap, psfs, dy, interpmeth=method, mode=mode)
    assert plan.info()["conv"] == ("separable" if separable else "direct")
    ang = angles32(nview).astype(np.float64)
    P32 = O.Problem(mumap, psfs, dy, ang, method, np.float32, "fft")
    P64 = O.Problem(mumap, psfs, dy, ang, method, np.float64, "direct")
    g = gpu.project(image, plan)
    r = P32.project(image)
    r64 = P64.project(image.astype(np.float64))
    assert per_view_maxrel(g, r) <= PROJ_TOL, per_view_maxrel(g, r)
    assert per_view_maxrel(g, r64) <= PROJ_TOL
    rng = np.random.default_rng(5)
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    gb = gpu.backproject(views, plan)
    rb = P32.backproject(views)
    assert maxrel(gb, rb) <= PROJ_TOL, maxrel(gb, rb)
    # subsets of views, in arbitrary order
    idx = [5, 0, 3]
    gs = gpu.project(image, plan, index=idx)
    assert per_view_maxrel(gs, g[:, :, idx]) <= 1e-6
    gbs = gpu.backproject(views[:, :, idx], plan, index=idx)
    rbs = P32.backproject(views[:, :, idx], index=np.array(idx, np.int32))
    assert maxrel(gbs, rbs) <= PROJ_TOL


@pytest.mark.parametrize("nx,nz,p", [(64, 40, 37), (64, 64, 37)])
def test_general_psf_large_planes(gpu, O, nx, nz, p):
    """Non-separable PSFs on planes whose adjoint staging takes a large part
    of LDS, so that the strip adjoint runs 512- and 1024-thread blocks.  Its
    staging holds the padded view (odd_up(nx + 36) x (nz + 36) floats), the
    strip windows (nz x odd_up(nx + 36 + 40 + 16)) and the taps (37 x 40):
    64x40 takes 61,744 B, so two blocks fit in 160 KiB and a block has 512
    threads; 64x64 takes 86,512 B, one block, 1024 threads.  Parity with the
    direct FP32 oracle, the FP64 oracle and the adjoint identity."""
    nview = 3
    rng = np.random.default_rng(nx + nz + p)
    mumap = np.asfortranarray((0.05 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = rand_sym_psfs(rng, p, p, nx, nview)
    plan = gpu.SPECTplan(mumap, psfs, 0.48, conv="direct")  # (the default takes the FFT path for 37 x 37)
    assert plan.info()["conv"] == "direct"
    ang = angles32(nview).astype(np.float64)
    P32 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float32, "direct")
    P64 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float64, "direct")
    image = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    g = gpu.project(image, plan)
    assert per_view_maxrel(g, P32.project(image)) <= PROJ_TOL
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    gb = gpu.backproject(views, plan)
    assert maxrel(gb, P32.backproject(views)) <= PROJ_TOL
    assert maxrel(gb, P64.backproject(views.astype(np.float64))) <= 1e-5
    a = float(np.vdot(views.astype(np.float64), g.astype(np.float64)))
    b = float(np.vdot(image.astype(np.float64), gb.astype(np.float64)))
    assert abs(a - b) <= 1e-5 * abs(a)


@pytest.mark.parametrize("method", ["one", "two"])
def test_adjoint_dot(gpu, method):
    """<y, Ax> = <A'y, x> at 32x32x20, 30 views, 5x3 PSF (test/adjoint-project.jl:52-107)."""
    nx, nz, nview = 32, 20, 30
    rng = np.random.default_rng(17)
    mumap = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    psfs = rand_sym_psfs(rng, 5, 3, nx, nview)
    plan = gpu.SPECTplan(mumap, psfs, 4.7952, interpmeth=method)
    x = np.asfortranarray(rng.standard_normal((nx, nx, nz)).astype(np.float32))
    y = np.asfortranarray(rng.standard_normal((nx, nz, nview)).astype(np.float32))
    Ax = gpu.project(x, plan)
    Aty = gpu.backproject(y, plan)
    a = float(np.vdot(y.astype(np.float64), Ax.astype(np.float64)))
    b = float(np.vdot(x.astype(np.float64), Aty.astype(np.float64)))
    assert abs(a - b) <= 1e-5 * abs(a), (a, b)


@pytest.mark.parametrize("method", ["one", "two"])
def test_adjoint_dense(gpu, method):
    """Matrix(A)' == Matrix(A') at 8x8x6, 7 views, 5x3 PSF (test/adjoint-project.jl:12-49)."""
    nx, nz, nview = 8, 6, 7
    rng = np.random.default_rng(23)
    mumap = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    psfs = rand_sym_psfs(rng, 5, 3, nx, nview)
    plan = gpu.SPECTplan(mumap, psfs, 4.7952, interpmeth=method)
    N, M = nx * nx * nz, nx * nz * nview
    A = np.zeros((M, N))
    for k in range(N):
        e = np.zeros(N, np.float32)
        e[k] = 1
        A[:, k] = gpu.project(e.reshape(nx, nx, nz, order="F"), plan).ravel(order="F")
    At = np.zeros((N, M))
    for k in range(M):
        e = np.zeros(M, np.float32)
        e[k] = 1
        At[:, k] = gpu.backproject(e.reshape(nx, nz, nview, order="F"), plan).ravel(order="F")
    assert np.allclose(A.T, At, rtol=1e-5, atol=1e-6 * np.abs(A).max())


# ---------------------------------------------------------------------------
# ML-EM / OS-EM
# ---------------------------------------------------------------------------
def em_problem(nx=32, nz=16, nview=24, px=5):
    import oracle as O

    rng = np.random.default_rng(2)
    x = np.zeros((nx, nx, nz), np.float32)
    x[nx // 4:2 * nx // 3, nx // 5:3 * nx // 5, 2 * nz // 6:3 * nz // 6] = 1
    x[2 * nx // 5:3 * nx // 5, nx // 5:2 * nx // 5, 4 * nz // 6:5 * nz // 6] = 2
    psf = O.psf_gauss(nx, px, px, 1.0, 4.0)
    psfs = np.asfortranarray(np.repeat(psf[:, :, :, None], nview, 3))
    mumap = np.asfortranarray((0.02 * rng.random((nx, nx, nz))).astype(np.float32))
    return np.asfortranarray(x), mumap, psfs


def test_mlem_osem_vs_oracle(gpu, O):
    nx, nz, nview = 32, 16, 24
    xtrue, mumap, psfs = em_problem(nx, nz, nview)
    dy = 0.8
    plan = gpu.SPECTplan(mumap, psfs, dy)
    ang = angles32(nview).astype(np.float64)
    P64 = O.Problem(mumap, psfs, dy, ang, "two", np.float64, "direct")
    ytrue = gpu.project(xtrue, plan)
    b = np.full_like(ytrue, 0.1 * ytrue.mean())
    y = gpu.poisson(20 * (ytrue + b), seed=4) / 20
    x0 = np.ones_like(xtrue)
    A = gpu.spect_operator(plan)
    xg = gpu.mlem(x0, y, b, A, niter=5)
    xr = P64.mlem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), 5)
    rel = np.linalg.norm(xg - xr) / np.linalg.norm(xr)
    assert rel <= 1e-4, rel
    xg2 = np.zeros_like(x0)
    gpu.mlem_(xg2, x0, y, b, A, niter=5)
    # xhat1 ≈ xhat2 (test/ml-os-em.jl:45): isapprox with rtol sqrt(eps(Float32))
    assert np.linalg.norm(xg - xg2) <= np.sqrt(np.finfo(np.float32).eps) * np.linalg.norm(xg)
    Ab = gpu.Ablock(plan, 4)
    og = gpu.osem(x0, y, b, Ab, niter=3)
    orr = P64.osem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), 4, 3)
    rel = np.linalg.norm(og - orr) / np.linalg.norm(orr)
    assert rel <= 1e-4, rel


def test_errors_on_device(gpu):
    mumap, psfs, image = small_problem(1)
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    with pytest.raises(gpu.SPECTError):
        gpu.project(image, plan, index=[0, 99])
    Ab = gpu.Ablock(plan, 2)
    y = np.ones((16, 8, 8), np.float32)
    with pytest.raises(gpu.SPECTError, match="need background > 0"):
        gpu.osem(np.ones_like(image), y, np.zeros_like(y), Ab, niter=1)
    with pytest.raises(gpu.SPECTError, match="nview must be divisible by nblocks!"):
        gpu.Ablock(plan, 3)


def test_background_check_device_arrays(gpu):
    """b > 0 (src/ml-os-em.jl:30, :106) is checked on the device for device
    arrays, before any update: the output stays untouched."""
    from spectrecon_amd import _lib as L

    mumap, psfs, image = small_problem(1)
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    nx, ny, nz = image.shape
    nview = psfs.shape[3]
    y = np.ones((nx, nz, nview), np.float32)
    x0 = np.ones_like(image)
    sentinel = np.full_like(image, 7.0)
    for bad in (0.0, -1.0, np.nan):
        b = np.ones_like(y)
        b[3, 2, nview - 1] = bad
        d = [gpu.DeviceArray.from_host(a) for a in (sentinel, x0, y, b)]
        rc = gpu.lib().sr_osem(plan.handle, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, 2, 1, L.SR_MEM_DEVICE)
        assert rc == L.SR_ERR_ARG
        assert gpu.lib().sr_last_error().decode() == "need background > 0"
        assert np.array_equal(d[0].to_host(), sentinel)
    b = np.ones_like(y)
    d = [gpu.DeviceArray.from_host(a) for a in (sentinel, x0, y, b)]
    L.check(gpu.lib().sr_osem(plan.handle, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, 2, 1, L.SR_MEM_DEVICE))
    assert np.all(np.isfinite(d[0].to_host()))


def test_memcpy_d2d(gpu):
    """sr_memcpy_d2d: the library's streaming copy (16-byte vectors) plus the
    unaligned remainder, bit-exact."""
    from spectrecon_amd import _lib as L

    rng = np.random.default_rng(5)
    for n in (1, 3, 4, 1000, 4099, 1 << 20):
        a = rng.standard_normal(n).astype(np.float32)
        src, dst = gpu.DeviceArray.from_host(a), gpu.DeviceArray((n,))
        L.check(gpu.lib().sr_memcpy_d2d(dst.ptr, src.ptr, a.nbytes))
        assert np.array_equal(dst.to_host(), a)
    # unaligned pointers take the runtime copy
    a = rng.standard_normal(1001).astype(np.float32)
    src, dst = gpu.DeviceArray.from_host(a), gpu.DeviceArray((1001,))
    L.check(gpu.lib().sr_memcpy_d2d(dst.ptr + 4, src.ptr + 4, 4 * 1000))
    assert np.array_equal(dst.to_host()[1:], a[1:])


@pytest.mark.parametrize("method", ["one", "two"])
def test_two_d_mode(gpu, O, method):
    """2-D mode: nz = 1 with pz = 1 PSFs (docs/lit/examples/5-2d.jl:36-59)."""
    nx, nz, nview, px = 64, 1, 60, 13
    rng = np.random.default_rng(31)
    mumap = np.zeros((nx, nx, nz), np.float32)
    mumap[16:48, 20:44, 0] = 0.1
    psf = np.asarray(O.psf_gauss(nx, px, 1, 1.0, 4.0), np.float32)
    psfs = np.asfortranarray(np.repeat(psf[:, :, :, None], nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, 1.0, interpmeth=method)
    assert plan.info()["conv"] == "separable"
    ang = angles32(nview).astype(np.float64)
    P32 = O.Problem(mumap, psfs, 1.0, ang, method, np.float32, "fft")
    image = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    g = gpu.project(image, plan)
    assert g.shape == (nx, 1, nview)
    assert per_view_maxrel(g, P32.project(image)) <= PROJ_TOL
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    assert maxrel(gpu.backproject(views, plan), P32.backproject(views)) <= PROJ_TOL
    a = float(np.vdot(views.astype(np.float64), g.astype(np.float64)))
    b = float(np.vdot(image.astype(np.float64), gpu.backproject(views, plan).astype(np.float64)))
    assert abs(a - b) <= 1e-5 * abs(a)


@pytest.mark.parametrize("method", ["one", "two"])
def test_per_view_psfs_and_angles(gpu, O, method):
    """Views with their own (separable) PSF stacks and a non-default,
    non-uniform viewangle list (src/spectplan.jl:45,47,63)."""
    nx, nz, nview, px = 32, 12, 10, 9
    rng = np.random.default_rng(37)
    mumap = np.asfortranarray((0.05 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = np.asfortranarray(np.stack(
        [np.asarray(O.psf_gauss(nx, px, px, 0.5 + 0.2 * v, 2.0 + 0.3 * v), np.float32) for v in range(nview)], 3))
    ang = np.sort(rng.random(nview) * 2 * np.pi).astype(np.float32)
    plan = gpu.SPECTplan(mumap, psfs, 0.48, viewangle=ang, interpmeth=method)
    assert plan.info()["conv"] == "separable"
    P32 = O.Problem(mumap, psfs, 0.48, ang.astype(np.float64), method, np.float32, "fft")
    image = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    assert per_view_maxrel(gpu.project(image, plan), P32.project(image)) <= PROJ_TOL
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    assert maxrel(gpu.backproject(views, plan), P32.backproject(views)) <= PROJ_TOL
    idx = [7, 2]
    rbs = P32.backproject(views[:, :, idx], index=np.array(idx, np.int32))
    assert maxrel(gpu.backproject(views[:, :, idx], plan, index=idx), rbs) <= PROJ_TOL


@pytest.mark.parametrize("layout", ["auto", "zyx"])
@pytest.mark.parametrize("nx,nz,px,pz", [(32, 16, 7, 7), (64, 32, 15, 15), (48, 16, 37, 37), (32, 32, 9, 5),
                                          (16, 16, 3, 21), (32, 48, 5, 9)])
def test_project_matrix_core_sizes(gpu, O, monkeypatch, nx, nz, px, pz, layout):
    """Planes whose sides are multiples of 16 take the matrix-core separable
    PSF kernels, and (:two) the z-fastest work layout unless SR_LAYOUT=zyx;
    parity as for every other path, edge tiles included (PSF half-widths up
    to beyond one tile), and the adjoint identity."""
    if layout == "zyx":
        monkeypatch.setenv("SR_LAYOUT", "zyx")
    nview = 6
    rng = np.random.default_rng(nx + nz + px)
    mumap = np.asfortranarray((0.05 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = np.asfortranarray(np.repeat(np.asarray(O.psf_gauss(nx, px, pz, 1.0, 5.0), np.float32)[:, :, :, None],
                                       nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    assert plan.info()["conv"] == "separable"
    assert plan.info()["layout"] == ("z-fastest" if layout == "auto" else "x-fastest")
    ang = angles32(nview).astype(np.float64)
    P32 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float32, "fft")
    image = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    g = gpu.project(image, plan)
    assert per_view_maxrel(g, P32.project(image)) <= PROJ_TOL
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    gb = gpu.backproject(views, plan)
    assert maxrel(gb, P32.backproject(views)) <= PROJ_TOL
    idx = [4, 1]
    rbs = P32.backproject(views[:, :, idx], index=np.array(idx, np.int32))
    assert maxrel(gpu.backproject(views[:, :, idx], plan, index=idx), rbs) <= PROJ_TOL
    a = float(np.vdot(views.astype(np.float64), g.astype(np.float64)))
    b = float(np.vdot(image.astype(np.float64), gb.astype(np.float64)))
    assert abs(a - b) <= 1e-5 * abs(a)


def test_project_deep_planes(gpu, O):
    """Deep planes on the z-fastest path (80 x 128, 37-tap PSF): the 8-wave
    column adjoint (12 wave buffers do not fit in LDS here) and the
    4-producer forward with two gather rounds per plane; parity with the
    oracle and the adjoint identity.  The FP32 reference here is the oracle's
    direct (closed-form) convolution: its FFT form's own rounding reaches
    2.7e-5 of the FP64 back-projection at this plane's replicate-pad folds,
    where the device is within 2.2e-6 of FP64."""
    nx, nz, px, nview = 80, 128, 37, 4
    rng = np.random.default_rng(nz)
    mumap = np.asfortranarray((0.05 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = np.asfortranarray(np.repeat(np.asarray(O.psf_gauss(nx, px, px, 1.5, 6.0), np.float32)[:, :, :, None],
                                       nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    assert plan.info()["layout"] == "z-fastest"
    ang = angles32(nview).astype(np.float64)
    P32 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float32, "direct")
    P64 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float64, "direct")
    image = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    g = gpu.project(image, plan)
    assert per_view_maxrel(g, P32.project(image)) <= PROJ_TOL
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    gb = gpu.backproject(views, plan)
    assert maxrel(gb, P32.backproject(views)) <= PROJ_TOL
    assert maxrel(gb, P64.backproject(views.astype(np.float64))) <= 1e-5
    a = float(np.vdot(views.astype(np.float64), g.astype(np.float64)))
    b = float(np.vdot(image.astype(np.float64), gb.astype(np.float64)))
    assert abs(a - b) <= 1e-5 * abs(a)


def test_batches_z_fastest(gpu, O):
    """Launch batches smaller than the view count on a z-fastest plan: every
    batch staged separately, partial images accumulated, the EM update on
    the multi-batch path; equal to the one-batch results and to the oracle."""
    nx, nz, nview = 32, 16, 12
    xtrue, mumap, psfs = em_problem(nx, nz, nview)
    plan = gpu.SPECTplan(mumap, psfs, 0.8)
    assert plan.info()["layout"] == "z-fastest"
    P32 = O.Problem(mumap, psfs, 0.8, angles32(nview).astype(np.float64), "two", np.float32, "fft")
    rng = np.random.default_rng(41)
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    g1, b1 = gpu.project(xtrue, plan), gpu.backproject(views, plan)
    ytrue = g1
    b = np.full_like(ytrue, 0.1 * ytrue.mean())
    y = gpu.poisson(20 * (ytrue + b), seed=9) / 20
    x0 = np.ones_like(xtrue)
    m1 = gpu.mlem(x0, y, b, gpu.spect_operator(plan), niter=3)
    o1 = gpu.osem(x0, y, b, gpu.Ablock(plan, 2), niter=2)
    plan.set_batch(5)
    g5, b5 = gpu.project(xtrue, plan), gpu.backproject(views, plan)
    assert per_view_maxrel(g5, g1) <= 1e-6
    assert maxrel(b5, b1) <= 1e-6
    assert per_view_maxrel(g5, P32.project(xtrue)) <= PROJ_TOL
    assert maxrel(b5, P32.backproject(views)) <= PROJ_TOL
    m5 = gpu.mlem(x0, y, b, gpu.spect_operator(plan), niter=3)
    o5 = gpu.osem(x0, y, b, gpu.Ablock(plan, 2), niter=2)
    assert np.linalg.norm(m5 - m1) <= 1e-5 * np.linalg.norm(m1)
    assert np.linalg.norm(o5 - o1) <= 1e-5 * np.linalg.norm(o1)


def test_osem_per_view_psfs_z_fastest(gpu, O):
    """OS-EM on a z-fastest plan whose views carry their own PSF stacks: the
    normaliser back-projects ones view by view (no shared-PSF shortcut)."""
    nx, nz, nview, px = 32, 16, 8, 7
    xtrue, mumap, _ = em_problem(nx, nz, nview)
    psfs = np.asfortranarray(np.stack(
        [np.asarray(O.psf_gauss(nx, px, px, 0.8 + 0.1 * v, 3.0 + 0.2 * v), np.float32) for v in range(nview)], 3))
    plan = gpu.SPECTplan(mumap, psfs, 0.8)
    assert plan.info()["layout"] == "z-fastest"
    P64 = O.Problem(mumap, psfs, 0.8, angles32(nview).astype(np.float64), "two", np.float64, "direct")
    ytrue = gpu.project(xtrue, plan)
    b = np.full_like(ytrue, 0.1 * ytrue.mean())
    y = gpu.poisson(20 * (ytrue + b), seed=3) / 20
    x0 = np.ones_like(xtrue)
    og = gpu.osem(x0, y, b, gpu.Ablock(plan, 4), niter=2)
    orr = P64.osem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), 4, 2)
    assert np.linalg.norm(og - orr) <= 1e-4 * np.linalg.norm(orr)


def test_config0_mlem(gpu, O):
    """The small workload: 64x64x32 seeded phantom and mu-map, 60 views,
    15x15 Gaussian PSF (FWHM 1 -> 4 px over the 64 depths), dy = 0.48,
    y ~ Poisson(A x + 1), b = 1, 10 ML-EM iterations from ones; relative L2
    vs the FP64 oracle."""
    nx, nz, nview = 64, 32, 60
    image = gpu.phantom(nx, nx, nz, seed=0)
    mumap = gpu.mumap_phantom(nx, nx, nz)
    psf = gpu.psf_gauss(ny=nx, px=15, fwhm_start=1.0, fwhm_end=4.0)
    psfs = np.asfortranarray(np.repeat(psf[:, :, :, None], nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    ytrue = gpu.project(image, plan)
    b = np.ones_like(ytrue)
    y = gpu.poisson(ytrue + b, seed=1)
    x0 = np.ones_like(image)
    xg = gpu.mlem(x0, y, b, gpu.spect_operator(plan), niter=10)
    P64 = O.Problem(mumap, psfs, 0.48, gpu.default_viewangle(nview).astype(np.float64), "two", np.float64, "direct",
                    nthread=8)
    xr = P64.mlem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), 10)
    rel = np.linalg.norm(xg - xr) / np.linalg.norm(xr)
    assert rel <= 1e-4, rel


def test_config1_adjoint_identity(gpu):
    """<y, A x> = <A'y, x> at the full clinical size: 128x128x80, 128 views,
    37x37 depth-varying PSF (FWHM 1.5 -> 6 px), the benchmark's plan."""
    nx, nz, nview = 128, 80, 128
    mumap = gpu.mumap_phantom(nx, nx, nz)
    psf = gpu.psf_gauss(ny=nx, px=37, fwhm_start=1.5, fwhm_end=6.0)
    psfs = np.asfortranarray(np.repeat(psf[:, :, :, None], nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    assert plan.info()["layout"] == "z-fastest"
    rng = np.random.default_rng(57)
    x = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    yv = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    a = float(np.vdot(yv.astype(np.float64), gpu.project(x, plan).astype(np.float64)))
    b = float(np.vdot(x.astype(np.float64), gpu.backproject(yv, plan).astype(np.float64)))
    assert abs(a - b) <= 1e-5 * abs(a), (a, b)


def test_config1_benchmark_kernels_vs_oracle(gpu, O):
    """The benchmarked plan against the oracle: 128x128x80, 128 views, 37x37
    depth-varying Gaussian PSF (FWHM 1.5 -> 6 px), seeded phantom and mu-map,
    dy = 0.48, autotuned for 16-view subsets.  Projections and
    back-projections run one whole OS-EM subset (views 0, 8, ..., 120: view 0
    is theta = 0, views 32, 64 and 96 are quarter turns, the others lie in
    between), so the kernels run at the benchmark's launch shape: the fused
    rotation + PSF forward with all 8 consumer columns busy and 5 z tiles per
    wave, 16 plane chunks per view, the column adjoint and the 16-view list
    gather.  Bars: per-view max-relative <= 2e-5 vs the FP32 oracle (the
    reference's FFT path); <= 1e-5 vs the FP64 oracle; OS-EM 8 x 4 from y ~
    Poisson(s A x + b) (5e6 counts, b = 0.1 mean) within 1e-4 relative L2 of
    the FP64 oracle's OS-EM, which also runs the shared-PSF normaliser gather."""
    nx, nz, nview, nblocks, niter, dy = 128, 80, 128, 8, 4, 0.48
    image = gpu.phantom(nx, nx, nz, seed=0)
    mumap = gpu.mumap_phantom(nx, nx, nz)
    psf = gpu.psf_gauss(ny=nx, px=37, fwhm_start=1.5, fwhm_end=6.0)
    psfs = np.asfortranarray(np.repeat(psf[:, :, :, None], nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, dy)
    assert plan.info()["layout"] == "z-fastest"
    plan.autotune(nview // nblocks)
    ang = gpu.default_viewangle(nview).astype(np.float64)
    P32 = O.Problem(mumap, psfs, dy, ang, "two", np.float32, "fft", nthread=16)
    P64 = O.Problem(mumap, psfs, dy, ang, "two", np.float64, "fft", nthread=16)
    for first in (0, 3):
        idx = np.arange(first, nview, nblocks, dtype=np.int32)
        g = gpu.project(image, plan, index=idx)
        e32 = per_view_maxrel(g, P32.project(image, idx))
        e64 = per_view_maxrel(g, P64.project(image.astype(np.float64), idx))
        print(f"subset {first}: project vs FP32 {e32:.2e}, vs FP64 {e64:.2e}", flush=True)
        assert e32 <= PROJ_TOL and e64 <= 1e-5, (first, e32, e64)
        rng = np.random.default_rng(first)
        views = np.asfortranarray((g * (0.5 + rng.random(g.shape))).astype(np.float32))
        gb = gpu.backproject(views, plan, index=idx)
        b32 = maxrel(gb, P32.backproject(views, idx))
        b64 = maxrel(gb, P64.backproject(views.astype(np.float64), idx))
        print(f"subset {first}: backproject vs FP32 {b32:.2e}, vs FP64 {b64:.2e}", flush=True)
        assert b32 <= PROJ_TOL and b64 <= 1e-5, (first, b32, b64)
    ytrue = gpu.project(image, plan)
    s = 5e6 / float(ytrue.sum())
    b = np.full_like(ytrue, 0.1 * float((s * ytrue).mean()))
    y = gpu.poisson(s * ytrue + b, seed=0)
    x0 = np.ones_like(image)
    xg = gpu.osem(x0, y, b, gpu.Ablock(plan, nblocks), niter=niter)
    xr = P64.osem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), nblocks, niter)
    rel = float(np.linalg.norm(xg - xr) / np.linalg.norm(xr))
    print(f"OS-EM 8x4 vs FP64: relative L2 {rel:.2e}", flush=True)
    assert rel <= 1e-4, rel


@pytest.mark.parametrize("nview", [16, 6])
def test_gradient_vs_oracle(gpu, O, nview):
    """L = ||A(x, mu) - y||^2 and its gradients w.r.t. x and the mu-map at
    32x32x16 (the differentiable-projector check size) against the FP64
    oracle, whose gradients are pinned by finite differences; all views and a
    view subset."""
    nx, nz, dy = 32, 16, 0.48
    rng = np.random.default_rng(71)
    mumap = np.asfortranarray((0.2 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = np.asfortranarray(np.repeat(np.asarray(O.psf_gauss(nx, 7, 7, 1.0, 3.0), np.float32)[:, :, :, None],
                                       nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, dy)
    assert plan.info()["layout"] == "z-fastest"
    x = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    y = np.asfortranarray((0.5 * gpu.project(np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32)), plan))
                          .astype(np.float32))
    P64 = O.Problem(mumap, psfs, dy, angles32(nview).astype(np.float64), "two", np.float64, "direct")
    for idx in (None, [nview - 1, 0, 3]):
        yi = y if idx is None else np.asfortranarray(y[:, :, idx])
        loss, gx, gmu = gpu.gradient(x, yi, plan, index=idx)
        lr, gxr, gmr = P64.grad(x.astype(np.float64), yi.astype(np.float64),
                                None if idx is None else np.array(idx, np.int32))
        assert abs(loss - lr) <= 1e-5 * lr, (loss, lr)
        assert np.linalg.norm(gx - gxr) <= 1e-4 * np.linalg.norm(gxr)
        assert np.linalg.norm(gmu - gmr) <= 1e-4 * np.linalg.norm(gmr), np.linalg.norm(gmu - gmr) / np.linalg.norm(gmr)


@pytest.mark.parametrize("nview", [16, 6])
def test_project_vjp_mu_vs_oracle(gpu, O, nview):
    """The mu vector-Jacobian product d<dy, A(x, mu)>/dmu against the FP64
    oracle's mu-gradient at y = A x - dy/2, whose residual 2 (A x - y) is dy
    (the oracle's gradient is pinned by finite differences); all views and a
    view subset; the image part of the product is the back-projection."""
    nx, nz, dy_cm = 32, 16, 0.48
    rng = np.random.default_rng(73)
    mumap = np.asfortranarray((0.2 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = np.asfortranarray(np.repeat(np.asarray(O.psf_gauss(nx, 7, 7, 1.0, 3.0), np.float32)[:, :, :, None],
                                       nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, dy_cm)
    x = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    P64 = O.Problem(mumap, psfs, dy_cm, angles32(nview).astype(np.float64), "two", np.float64, "direct")
    for idx in (None, [nview - 1, 0, 3]):
        n = nview if idx is None else len(idx)
        cot = np.asfortranarray(rng.standard_normal((nx, nz, n)).astype(np.float32))
        gmu = gpu.project_vjp_mu(x, cot, plan, index=idx)
        ia = None if idx is None else np.array(idx, np.int32)
        y64 = P64.project(x.astype(np.float64), ia) - 0.5 * cot.astype(np.float64)
        _, _, gmr = P64.grad(x.astype(np.float64), y64, ia)
        assert np.linalg.norm(gmu - gmr) <= 1e-4 * np.linalg.norm(gmr), np.linalg.norm(gmu - gmr) / np.linalg.norm(gmr)
    with pytest.raises(gpu.DimensionMismatch):
        gpu.project_vjp_mu(x, cot[:, :, :1], plan, index=[0, 1])


def test_named_adjoint_entry_points(gpu):
    """sr_imrotate_adj and sr_fft_conv_adj are sr_imrotate / sr_fft_conv with adj = 1."""
    from spectrecon_amd import _lib as L
    rng = np.random.default_rng(5)
    img = np.asfortranarray(rng.random((24, 24, 3), dtype=np.float32))
    a, b = np.zeros_like(img), np.zeros_like(img)
    for interp in (L.SR_INTERP_ONE, L.SR_INTERP_TWO):
        L.check(gpu.lib().sr_imrotate(a.ctypes.data, img.ctypes.data, 24, 3, 0.7, interp, 1, L.SR_MEM_HOST))
        L.check(gpu.lib().sr_imrotate_adj(b.ctypes.data, img.ctypes.data, 24, 3, 0.7, interp, L.SR_MEM_HOST))
        np.testing.assert_array_equal(a, b)
    im2 = np.asfortranarray(rng.random((20, 14), dtype=np.float32))
    ker = np.asfortranarray(np.outer(np.hanning(7), np.hanning(5)).astype(np.float32))
    c, d = np.zeros_like(im2), np.zeros_like(im2)
    L.check(gpu.lib().sr_fft_conv(c.ctypes.data, im2.ctypes.data, 20, 14, ker.ctypes.data, 7, 5, 1, L.SR_MEM_HOST))
    L.check(gpu.lib().sr_fft_conv_adj(d.ctypes.data, im2.ctypes.data, 20, 14, ker.ctypes.data, 7, 5, L.SR_MEM_HOST))
    np.testing.assert_array_equal(c, d)
    assert np.abs(c).sum() > 0


@pytest.mark.parametrize("nview", [40, 7])
def test_gradient_equals_operators(gpu, O, nview):
    """sr_grad takes both gradients from one PSF adjoint per launch batch (40
    views: two batches); they equal the separate operators bit for bit:
    dL/dx = backproject(r) and dL/dmu = project_vjp_mu(x, r) with
    r = 2 (A x - y), for all views and a view subset."""
    nx, nz, dy = 32, 16, 0.48
    rng = np.random.default_rng(97)
    mumap = np.asfortranarray((0.2 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = np.asfortranarray(np.repeat(np.asarray(O.psf_gauss(nx, 7, 7, 1.0, 3.0), np.float32)[:, :, :, None],
                                       nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, dy)
    assert plan.info()["layout"] == "z-fastest"
    x = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    y = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    for idx in (None, [nview - 1, 0, 3, 2]):
        yi = y if idx is None else np.asfortranarray(y[:, :, idx])
        loss, gx, gmu = gpu.gradient(x, yi, plan, index=idx)
        r = np.asfortranarray((2 * (gpu.project(x, plan, index=idx) - yi)).astype(np.float32))
        np.testing.assert_array_equal(gx, gpu.backproject(r, plan, index=idx))
        np.testing.assert_array_equal(gmu, gpu.project_vjp_mu(x, r, plan, index=idx))
        assert abs(loss - 0.25 * float(np.sum(r.astype(np.float64) ** 2))) <= 1e-6 * loss


def test_gradient_empty_index(gpu, O):
    """sr_grad over no view: the loss is 0 and both gradients are zero."""
    nx, nz, nview = 32, 16, 4
    rng = np.random.default_rng(3)
    mumap = np.asfortranarray((0.1 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = np.asfortranarray(np.repeat(np.asarray(O.psf_gauss(nx, 5, 5, 1.0, 2.0), np.float32)[:, :, :, None],
                                       nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    x = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    gpu.gradient(x, np.asfortranarray(np.ones((nx, nz, nview), np.float32)), plan)  # leaves a loss behind
    loss, gx, gmu = gpu.gradient(x, np.zeros((nx, nz, 0), np.float32), plan, index=[])
    assert loss == 0.0
    assert not gx.any() and not gmu.any()


def test_gradient_needs_z_fastest(gpu, monkeypatch):
    monkeypatch.setenv("SR_LAYOUT", "zyx")
    mumap, psfs, image = small_problem(3, 32, 16, 4, 5, 5, True)
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    with pytest.raises(gpu.DeviceError, match="z-fastest"):
        gpu.gradient(image, np.zeros((32, 16, 4), np.float32), plan)


def test_osem_strided_uneven_vs_oracle(gpu, O):
    """OS-EM over strided subsets whose sizes differ (10 views, 4 blocks:
    3, 3, 2, 2 views) against the FP64 oracle; with nblocks dividing nview it
    is osem itself."""
    nx, nz, nview = 32, 16, 10
    xtrue, mumap, psfs = em_problem(nx, nz, nview)
    plan = gpu.SPECTplan(mumap, psfs, 0.8)
    ytrue = gpu.project(xtrue, plan)
    b = np.full_like(ytrue, 0.1 * ytrue.mean())
    y = gpu.poisson(20 * (ytrue + b), seed=5) / 20
    x0 = np.ones_like(xtrue)
    P64 = O.Problem(mumap, psfs, 0.8, angles32(nview).astype(np.float64), "two", np.float64, "direct")
    xg = gpu.osem_strided(x0, y, b, plan, 4, 2)
    xr = P64.osem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), 4, 2)
    assert np.linalg.norm(xg - xr) <= 1e-4 * np.linalg.norm(xr)
    with pytest.raises(gpu.SPECTError, match="nview must be divisible by nblocks!"):
        gpu.Ablock(plan, 4)
    np.testing.assert_array_equal(gpu.osem_strided(x0, y, b, plan, 5, 2),
                                  gpu.osem(x0, y, b, gpu.Ablock(plan, 5), niter=2))


@pytest.mark.parametrize("nx,nz,nview,nblocks", [(32, 16, 16, 4), (32, 16, 10, 4), (64, 32, 60, 8)])
def test_em_normaliser_placements(gpu, O, nx, nz, nview, nblocks):
    """OS-EM on a z-fastest plan with one PSF stack for all views sums its
    subset normalisers A_b' 1 either in each subset's first-iteration gather
    (em_norm 1, the built-in) or all in one gather launch before the
    iterations (em_norm 2, grid row k = block k, the first nview % nblocks
    blocks one view longer).
    Both against the FP64 oracle, even and uneven strided blocks, and each
    normaliser placement against the other."""
    xtrue, mumap, psfs = em_problem(nx, nz, nview)
    plan = gpu.SPECTplan(mumap, psfs, 0.8)
    assert plan.info()["layout"] == "z-fastest"
    ytrue = gpu.project(xtrue, plan)
    b = np.full_like(ytrue, 0.1 * ytrue.mean())
    y = gpu.poisson(20 * (ytrue + b), seed=7) / 20
    x0 = np.ones_like(xtrue)
    P64 = O.Problem(mumap, psfs, 0.8, angles32(nview).astype(np.float64), "two", np.float64, "direct")
    xr = P64.osem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), nblocks, 2)
    got = {}
    for m in (1, 2, 0):
        plan.set_choice(em_norm=m)
        got[m] = gpu.osem_strided(x0, y, b, plan, nblocks, 2)
        err = np.linalg.norm(got[m] - xr) / np.linalg.norm(xr)
        assert err <= 1e-4, (m, err)
    assert np.linalg.norm(got[1] - got[2]) <= 1e-6 * np.linalg.norm(got[2])
    np.testing.assert_array_equal(got[0], got[1])
    with pytest.raises(gpu.SPECTError, match="EM normaliser choice"):
        plan.set_choice(em_norm=3)


CHOICE_CASES = {
    "tile_kernels": dict(psf_fwd_family=1, psf_adj_family=1),
    "column_fwd": dict(psf_fwd_family=2, psf_adj_family=2, gather_slices=8),
    "overlap8_fwd": dict(psf_fwd_family=3, psf_fwd_chunk_percent=200),
    "overlap4_fwd": dict(psf_fwd_family=4, psf_fwd_chunk_percent=50, gather_slices=2),
    "pairs_fwd": dict(psf_fwd_family=5),
    "pairs_fwd_chunks": dict(psf_fwd_family=5, psf_fwd_chunk_percent=400),
    "half_chunks": dict(psf_fwd_chunk_percent=50, psf_adj_chunk_percent=50, gather_slices=2),
    "quad_chunks": dict(psf_fwd_chunk_percent=400, psf_adj_chunk_percent=400, gather_slices=4),
    "tile_double": dict(psf_fwd_family=1, psf_adj_family=1, psf_fwd_chunk_percent=200,
                        psf_adj_chunk_percent=200, gather_slices=8),
}


@pytest.mark.parametrize("case", sorted(CHOICE_CASES))
@pytest.mark.parametrize("nx,nz,px,pz", [(32, 16, 7, 7), (64, 32, 15, 15), (48, 16, 37, 37), (16, 16, 3, 21),
                                          (32, 48, 5, 9)])
def test_kernel_choices_vs_oracle(gpu, O, nx, nz, px, pz, case):
    """Every kernel choice the autotuner can make (PSF kernel family, plane
    chunks per view, gather view slices) computes the same operator: parity
    with the oracle as for the built-in choices, and the adjoint identity."""
    nview = 6
    rng = np.random.default_rng(nx + nz + px)
    mumap = np.asfortranarray((0.05 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = np.asfortranarray(np.repeat(np.asarray(O.psf_gauss(nx, px, pz, 1.0, 5.0), np.float32)[:, :, :, None],
                                       nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, 0.48)
    assert plan.info()["layout"] == "z-fastest"
    plan.set_choice(**CHOICE_CASES[case])
    assert {k: plan.choice()[k] for k in CHOICE_CASES[case]} == CHOICE_CASES[case]
    P32 = O.Problem(mumap, psfs, 0.48, angles32(nview).astype(np.float64), "two", np.float32, "fft")
    image = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    g = gpu.project(image, plan)
    assert per_view_maxrel(g, P32.project(image)) <= PROJ_TOL
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    gb = gpu.backproject(views, plan)
    assert maxrel(gb, P32.backproject(views)) <= PROJ_TOL
    a = float(np.vdot(views.astype(np.float64), g.astype(np.float64)))
    b = float(np.vdot(image.astype(np.float64), gb.astype(np.float64)))
    assert abs(a - b) <= 1e-5 * abs(a)


def test_autotune_records_and_keeps_choices(gpu, O):
    """The autotuner times every candidate, keeps the fastest of each choice,
    logs them, and the tuned plan's OS-EM matches the FP64 oracle; invalid
    choices are rejected and nsub <= 0 restores the built-in ones."""
    nx, nz, nview = 32, 16, 12
    xtrue, mumap, psfs = em_problem(nx, nz, nview)
    plan = gpu.SPECTplan(mumap, psfs, 0.8)
    rec = plan.autotune(3)
    assert rec["views_per_launch"] == 3 and rec["layout"] == "z-fastest"
    ch = plan.choice()
    for key, fam, pct in (("psf_fwd", "psf_fwd_family", "psf_fwd_chunk_percent"),
                          ("psf_adj", "psf_adj_family", "psf_adj_chunk_percent")):
        cands = rec[key]["candidates"]
        assert len(cands) >= 2 and all(c["us"] > 0 for c in cands)
        best = min(cands, key=lambda c: c["us"])
        # the fastest, unless the built-in choice is within the tuning margin
        builtin = rec[key]["choice"]["code"] == 0 and rec[key]["choice"]["chunk_percent"] == 100
        assert rec[key]["choice"]["us"] == best["us"] or (builtin and best["us"] > 0.97 * rec[key]["choice"]["us"])
        assert ch[fam] == rec[key]["choice"]["code"]
        assert rec[key]["choice"]["family"] in ("overlap", "overlap8", "column", "tile")
        assert ch[pct] == rec[key]["choice"]["chunk_percent"]
    assert [c["slices"] for c in rec["gather"]["candidates"]] == [0, 2, 4, 8]
    assert ch["gather_slices"] == rec["gather"]["choice"]["slices"]
    ytrue = gpu.project(xtrue, plan)
    b = np.full_like(ytrue, 0.1 * ytrue.mean())
    y = gpu.poisson(20 * (ytrue + b), seed=5) / 20
    x0 = np.ones_like(xtrue)
    P64 = O.Problem(mumap, psfs, 0.8, angles32(nview).astype(np.float64), "two", np.float64, "direct")
    xg = gpu.osem_strided(x0, y, b, plan, 4, 2)
    xr = P64.osem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), 4, 2)
    assert np.linalg.norm(xg - xr) <= 1e-4 * np.linalg.norm(xr)
    assert {c["family"] for c in rec["psf_fwd"]["candidates"]} == {"overlap", "overlap8", "pairs", "column", "tile"}
    for bad in (dict(psf_fwd_family=6), dict(psf_adj_family=3), dict(psf_adj_chunk_percent=5),
                dict(gather_slices=3)):
        with pytest.raises(gpu.SPECTError):
            plan.set_choice(**bad)
    assert plan.autotune(0) == {"reset": True}
    assert list(plan.choice().values()) == [0, 100, 0, 100, 0, 0, 0]


def test_autotune_other_plans(gpu, O, monkeypatch):
    """x-fastest plans tune only the PSF chunk counts (one family, no gather);
    plans with a general (non-separable) PSF have nothing to tune; both still
    project as the oracle does afterwards."""
    nx, nz, nview = 32, 16, 8
    xtrue, mumap, psfs = em_problem(nx, nz, nview)
    monkeypatch.setenv("SR_LAYOUT", "zyx")
    plan = gpu.SPECTplan(mumap, psfs, 0.8)
    assert plan.info()["layout"] == "x-fastest"
    rec = plan.autotune(4)
    assert rec["layout"] == "x-fastest" and "gather" not in rec
    assert {c["family"] for c in rec["psf_fwd"]["candidates"]} == {"tile"}
    P32 = O.Problem(mumap, psfs, 0.8, angles32(nview).astype(np.float64), "two", np.float32, "fft")
    assert per_view_maxrel(gpu.project(xtrue, plan), P32.project(xtrue)) <= PROJ_TOL
    monkeypatch.delenv("SR_LAYOUT")
    rng = np.random.default_rng(9)
    gpsf = rand_sym_psfs(rng, 5, 5, nx, nview)  # not rank 1: the direct 2-D kernels
    plan2 = gpu.SPECTplan(mumap, gpsf, 0.8)
    assert plan2.info()["conv"] == "direct"
    rec2 = plan2.autotune(4)
    assert "psf_fwd" not in rec2 and "gather" not in rec2
    P32b = O.Problem(mumap, gpsf, 0.8, angles32(nview).astype(np.float64), "two", np.float32, "fft")
    assert per_view_maxrel(gpu.project(xtrue, plan2), P32b.project(xtrue)) <= PROJ_TOL


@pytest.mark.parametrize("nx,nz,nview,angles", [
    (128, 80, 128, None),  # the benchmark geometry
    (96, 32, 60, None),
    (64, 16, 9, "mixed"),  # 0, quarter turns, their neighbours, arbitrary angles
    (16, 16, 8, "mixed"),  # the smallest z-fastest planes
    (160, 32, 6, None),    # the largest z-fastest planes in x (the tile-distributed PSF kernels)
])
def test_device_lists_equal_host_lists(gpu, O, monkeypatch, nx, nz, nview, angles):
    """The rotation-adjoint lists built on the device at plan creation equal
    the host builder's (SR_LISTS=host): back-projections and an OS-EM pass on
    the two plans agree bit for bit (same sources, weights and slot order);
    the device-built plan also matches the FP32 oracle's back-projection."""
    xtrue, mumap, psfs = em_problem(nx, nz, nview)
    ang = None
    if angles == "mixed":
        half = np.float32(np.pi / 2)
        ang = np.array([0, half, 2 * half, 3 * half, np.nextafter(half, np.float32(4)), 0.3, 1.0, 2.5, 5.9],
                       np.float32)[:nview]
    print(f"lists: {nx}x{nx}x{nz}, {nview} views, angles={angles}", flush=True)
    dev = gpu.SPECTplan(mumap, psfs, 0.8, viewangle=ang)
    monkeypatch.setenv("SR_LISTS", "host")
    host = gpu.SPECTplan(mumap, psfs, 0.8, viewangle=ang)
    monkeypatch.delenv("SR_LISTS")
    assert dev.info()["layout"] == "z-fastest" and host.info()["layout"] == "z-fastest"
    rng = np.random.default_rng(nx + nview)
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    bd = gpu.backproject(views, dev)
    np.testing.assert_array_equal(bd, gpu.backproject(views, host))
    y = gpu.project(xtrue, dev) + 0.1
    b = np.full_like(y, 0.1)
    x0 = np.ones_like(xtrue)
    nb = 3 if nview % 3 == 0 else 1
    np.testing.assert_array_equal(gpu.osem(x0, y, b, gpu.Ablock(dev, nb), niter=1),
                                  gpu.osem(x0, y, b, gpu.Ablock(host, nb), niter=1))
    a = angles32(nview).astype(np.float64) if ang is None else ang.astype(np.float64)
    P32 = O.Problem(mumap, psfs, 0.8, a, "two", np.float32, "fft", nthread=16)
    assert maxrel(bd, P32.backproject(views)) <= PROJ_TOL


@pytest.mark.parametrize("variant", ["shared", "per_view_psf", "x_fastest", "batch2"])
def test_graph_replay_equals_eager(gpu, O, monkeypatch, variant):
    """EM calls replayed from a HIP graph equal the eager calls bit for bit,
    on every replay, with other work of the plan (a projection that regrows
    its work buffers, so the graph is captured again) between the replays;
    device-mode replays read the current contents of their inputs.  Variants
    cover the four kernel sequences: z-fastest with one PSF (shared
    normaliser), per-view PSFs, the x-fastest layout, and subsets larger than
    the launch batch."""
    from spectrecon_amd import _lib as L
    nx, nz, nview = 32, 16, 10
    xtrue, mumap, psfs = em_problem(nx, nz, nview)
    if variant == "per_view_psf":
        psfs = np.asfortranarray(np.stack(
            [np.asarray(O.psf_gauss(nx, 5, 5, 0.8 + 0.1 * v, 3.0 + 0.2 * v), np.float32) for v in range(nview)], 3))
    if variant == "x_fastest":
        monkeypatch.setenv("SR_LAYOUT", "zyx")
    eager, graphed = gpu.SPECTplan(mumap, psfs, 0.8), gpu.SPECTplan(mumap, psfs, 0.8)
    if variant == "batch2":
        eager.set_batch(2)
        graphed.set_batch(2)
    graphed.set_graph(True)
    ytrue = gpu.project(xtrue, eager)
    b = np.full_like(ytrue, 0.1 * ytrue.mean())
    ys = [gpu.poisson(20 * (ytrue + b), seed=s) / 20 for s in (3, 4)]
    x0 = np.ones_like(xtrue)
    ref = [gpu.osem_strided(x0, y, b, eager, 4, 2) for y in ys]
    for k in range(3):
        np.testing.assert_array_equal(gpu.osem_strided(x0, ys[0], b, graphed, 4, 2), ref[0])
    gpu.project(xtrue, graphed)  # all views in one batch: the work buffers grow
    for k in range(3):
        np.testing.assert_array_equal(gpu.osem_strided(x0, ys[0], b, graphed, 4, 2), ref[0])
    cap, rep = graphed.graph_stats()
    # (captured once more if the projection moved a buffer, else replayed)
    assert (cap, rep) in ((2, 2), (1, 4)), (cap, rep)
    # device arrays: replays see new measurement contents
    dx, dy_, db, dout = (gpu.DeviceArray.from_host(a) for a in (x0, ys[0], b, x0))
    for k in range(4):
        y = ys[k % 2]
        dy_.copy_from(y)
        L.check(gpu.lib().sr_osem_strided(graphed.handle, dout.ptr, dx.ptr, dy_.ptr, db.ptr, 4, 2, L.SR_MEM_DEVICE))
        np.testing.assert_array_equal(dout.to_host(), ref[k % 2])
    assert graphed.graph_stats() == (cap + 1, rep + 2)
    graphed.set_graph(False)
    np.testing.assert_array_equal(gpu.osem_strided(x0, ys[1], b, graphed, 4, 2), ref[1])
    assert graphed.graph_stats() == (cap + 1, rep + 2)


def test_shared_launch_batch_async_mixed(gpu, O):
    """sr_osem_batch with SR_ASYNC over a mix: three z-fastest instances (even
    and uneven strided subsets; one with the 4-producer forward and pinned
    gather slices, so that it shares no launch with the same-size instance
    on the built-in choices) run as shared launches, an x-fastest one
    (non-separable PSFs) on its own stream; after sr_plan_synchronize per
    plan, the shared results are within 1e-5 relative L2 of each instance
    alone and identical on every run (eager, graph capture, replay), the
    x-fastest one equals its own call bit for bit, and all match the FP64
    oracle."""
    import ctypes as C

    from spectrecon_amd import _lib as L
    from spectrecon_amd import serve as V
    from spectrecon_amd.device import DeviceArray
    rng = np.random.default_rng(11)
    cases = []
    for nx, nz, nview, nblocks, general, choice in (
            (32, 16, 16, 4, False, {}),
            (64, 32, 12, 5, False, dict(psf_fwd_family=4, gather_slices=4)),  # pinned gather slices
            (64, 32, 12, 3, False, {}),
            (32, 16, 10, 3, True, {})):
        xtrue, mumap, psfs = em_problem(nx, nz, nview)
        if general:
            psfs = rand_sym_psfs(rng, 5, 5, nx, nview)
        plan = gpu.SPECTplan(mumap, psfs, 0.8)
        assert plan.info()["layout"] == ("x-fastest" if general else "z-fastest")
        plan.set_graph(True)
        if choice:
            plan.set_choice(**choice)
        ytrue = gpu.project(xtrue, plan)
        b = np.full_like(ytrue, 0.1 * ytrue.mean())
        y = gpu.poisson(20 * (ytrue + b), seed=nx + nview) / 20
        x0 = np.ones_like(xtrue)
        alone = gpu.osem_strided(x0, y, b, plan, nblocks, 2)
        P64 = O.Problem(mumap, psfs, 0.8, angles32(nview).astype(np.float64), "two", np.float64, "direct")
        xr = P64.osem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), nblocks, 2)
        dev = [DeviceArray.from_host(a) for a in (x0, y, b)] + [DeviceArray(x0.shape)]
        cases.append(dict(plan=plan, nblocks=nblocks, alone=alone, xr=xr, general=general, dev=dev))
    n = len(cases)
    P = (C.c_void_p * n)(*[c["plan"].handle for c in cases])
    arr = [(C.c_void_p * n)(*[c["dev"][k].ptr for c in cases]) for k in range(4)]
    nb = (C.c_int32 * n)(*[c["nblocks"] for c in cases])
    ni = (C.c_int32 * n)(*([2] * n))
    lib = L.lib()
    s0 = V.batch_stats()
    runs = []
    for rnd in range(3):
        L.check(lib.sr_osem_batch(n, P, arr[3], arr[0], arr[1], arr[2], nb, ni, L.SR_MEM_DEVICE | L.SR_ASYNC))
        for c in cases:
            c["plan"].synchronize()
        runs.append([c["dev"][3].to_host() for c in cases])
    s1 = V.batch_stats()
    assert s1["instances"] - s0["instances"] == 3 * 3
    assert (s1["captures"] - s0["captures"], s1["replays"] - s0["replays"]) == (1, 1)
    for k, c in enumerate(cases):
        for rnd in range(1, 3):
            np.testing.assert_array_equal(runs[rnd][k], runs[0][k], err_msg=f"instance {k}, run {rnd}")
        got = runs[0][k]
        if c["general"]:
            np.testing.assert_array_equal(got, c["alone"])
        else:
            assert np.linalg.norm(got - c["alone"]) <= 1e-5 * np.linalg.norm(c["alone"]), k
        assert np.linalg.norm(got - c["xr"]) <= 1e-4 * np.linalg.norm(c["xr"]), k
    for c in cases:
        for d in c["dev"]:
            d.free()
        c["plan"].close()


def test_serve_mix_equals_alone(gpu):
    """Requests of three sizes admitted concurrently by the serving loop (each
    on its own resident plan and stream): every result equals the same
    request reconstructed alone."""
    from spectrecon_amd import serve as V
    cfgs = [V.Config(32, 16, 16, 7, 0.48, 1.0, 3.0), V.Config(64, 32, 12, 15, 0.48, 1.0, 4.0),
            V.Config(48, 16, 10, 5, 0.48, 1.0, 2.0)]
    reqs = []
    for k in range(6):
        cfg = cfgs[k % 3]
        reqs.append(V.synthetic_request(cfg, seed=100 + k, nblocks=(4, 5, 3)[k % 3], niter=2, arrival=0.002 * k,
                                        rid=k, scale=1e5))
    server = V.Server(hbm_budget=8e9, max_active=4)
    try:
        st = server.run(reqs)
    finally:
        server.close()
    assert st["served"] == 6 and st["rejected"] == 0
    backend = V.DeviceBackend()
    for r in reqs:
        plan = backend.make_plan(r.config)
        try:
            alone = gpu.osem_strided(r.x0, r.y, r.b, plan, r.nblocks, r.niter)
        finally:
            plan.close()
        np.testing.assert_array_equal(r.result, alone)


@pytest.mark.parametrize("layout", ["z-fastest", "x-fastest"])
def test_set_mumap_equals_new_plan(gpu, O, monkeypatch, layout):
    """sr_plan_set_mumap rebuilds every view's attenuation factors in place:
    projections and an OS-EM after it equal those of a plan built with the
    new mu-map bit for bit, and match the oracle (whose attenuation sums each
    y anew, the reference's order) at the projector bar."""
    if layout == "x-fastest":
        monkeypatch.setenv("SR_LAYOUT", "zyx")
    nx, nz, nview = 32, 16, 8
    rng = np.random.default_rng(21)
    psfs = np.asfortranarray(np.repeat(np.asarray(O.psf_gauss(nx, 5, 5, 1.0, 2.5), np.float32)[:, :, :, None],
                                       nview, 3))
    mu1 = np.asfortranarray((0.1 * rng.random((nx, nx, nz))).astype(np.float32))
    mu2 = np.asfortranarray((0.3 * rng.random((nx, nx, nz))).astype(np.float32))
    plan = gpu.SPECTplan(mu1, psfs, 0.8)
    fresh = gpu.SPECTplan(mu2, psfs, 0.8)
    assert plan.info()["layout"] == layout
    image = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    before = gpu.project(image, plan)
    plan.set_mumap(mu2)
    g = gpu.project(image, plan)
    np.testing.assert_array_equal(g, gpu.project(image, fresh))
    assert not np.array_equal(g, before)
    P32 = O.Problem(mu2, psfs, 0.8, angles32(nview).astype(np.float64), "two", np.float32, "fft")
    assert per_view_maxrel(g, P32.project(image)) <= PROJ_TOL
    b = np.full_like(g, 0.1 * g.mean())
    y = gpu.poisson(20 * (g + b), seed=2) / 20
    x0 = np.ones_like(image)
    np.testing.assert_array_equal(gpu.osem(x0, y, b, gpu.Ablock(plan, 4), niter=2),
                                  gpu.osem(x0, y, b, gpu.Ablock(fresh, 4), niter=2))
    d = gpu.DeviceArray.from_host(mu1)
    plan.set_mumap(d)
    np.testing.assert_array_equal(gpu.project(image, plan), before)


# ---------------------------------------------------------------------------
# PSF through the plane FFTs (x-fastest plans)
# ---------------------------------------------------------------------------
FFT_CASES = [
    # nx, nz, p, per-view PSFs: padded planes L1 x L2
    (32, 16, 7, False),    # 64 x 32
    (48, 7, 9, False),     # 64 x 16, odd nz (the last row pair has one row)
    (32, 15, 5, True),     # 64 x 32, odd nz, a PSF stack per view
    (64, 40, 19, False),   # 128 x 64: the reference's timing problem (time/project.jl)
    (16, 24, 9, False),    # 32 x 32
    (96, 64, 25, False),   # 128 x 128
    (128, 80, 37, False),  # 256 x 128: the clinical size with a non-separable PSF
    (8, 8, 5, False),      # 16 x 16, the smallest planes (a column half is 4 columns)
    (48, 110, 25, True),   # 128 x 256: the forward's opaque-index form with 16-point z passes
]


@pytest.mark.parametrize("nx,nz,p,per_view", FFT_CASES)
def test_fft_psf_vs_oracle(gpu, O, nx, nz, p, per_view):
    """Plans that run the PSF through the plane FFTs (conv="fft", x-fastest
    layout) against the oracle's literal FFT path (FP32, the reference's
    arithmetic) per view at <= 2e-5, against the FP64 oracle, and the adjoint
    identity; random point-symmetric PSFs, all views and an index subset."""
    nview = 4
    rng = np.random.default_rng(nx * 7 + nz + p)
    mumap = np.asfortranarray((0.05 * rng.random((nx, nx, nz))).astype(np.float32))
    if per_view:
        psfs = rand_sym_psfs(rng, p, p, nx, nview)
    else:
        psfs = np.asfortranarray(np.repeat(rand_sym_psfs(rng, p, p, nx, 1), nview, 3))
    print(f"fft psf: {nx}x{nx}x{nz}, {p}x{p}, per_view={per_view}", flush=True)
    plan = gpu.SPECTplan(mumap, psfs, 0.48, conv="fft")
    info = plan.info()
    assert info["conv"] == "fft" and info["layout"] == "x-fastest"
    ang = angles32(nview).astype(np.float64)
    P32 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float32, "fft")
    P64 = O.Problem(mumap, psfs, 0.48, ang, "two", np.float64, "fft")
    image = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    g = gpu.project(image, plan)
    assert per_view_maxrel(g, P32.project(image)) <= PROJ_TOL
    assert per_view_maxrel(g, P64.project(image.astype(np.float64))) <= 1e-5
    views = np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32))
    gb = gpu.backproject(views, plan)
    assert maxrel(gb, P32.backproject(views)) <= PROJ_TOL
    assert maxrel(gb, P64.backproject(views.astype(np.float64))) <= 1e-5
    a = float(np.vdot(views.astype(np.float64), g.astype(np.float64)))
    b = float(np.vdot(image.astype(np.float64), gb.astype(np.float64)))
    assert abs(a - b) <= 1e-5 * abs(a)
    idx = [3, 0]
    gi = gpu.project(image, plan, index=idx)
    assert per_view_maxrel(gi, P32.project(image, np.array(idx, np.int32))) <= PROJ_TOL
    gbi = gpu.backproject(np.asfortranarray(views[:, :, idx]), plan, index=idx)
    assert maxrel(gbi, P32.backproject(np.asfortranarray(views[:, :, idx]), np.array(idx, np.int32))) <= PROJ_TOL


@pytest.mark.parametrize("method", ["one", "two"])
def test_fft_psf_osem_and_methods(gpu, O, method):
    """OS-EM on a plan whose PSF runs through the plane FFTs, against the FP64
    oracle; the direct taps and the FFTs (set_choice psf_method) give the
    same operator; the autotuner times both and logs them."""
    nx, nz, nview, p = 32, 16, 8, 9
    rng = np.random.default_rng(5)
    mumap = np.asfortranarray((0.05 * rng.random((nx, nx, nz))).astype(np.float32))
    psfs = np.asfortranarray(np.repeat(rand_sym_psfs(rng, 11, 11, nx, 1), nview, 3))
    plan = gpu.SPECTplan(mumap, psfs, 0.8, interpmeth=method)
    assert plan.info()["conv"] == "fft"  # (11 x 11 general PSF: the FFTs are built in)
    xtrue = gpu.phantom(nx, nx, nz, seed=3)
    ytrue = gpu.project(xtrue, plan)
    b = np.full_like(ytrue, 0.1 * ytrue.mean())
    y = gpu.poisson(20 * (ytrue + b), seed=4) / 20
    x0 = np.ones_like(xtrue)
    xg = gpu.osem(x0, y, b, gpu.Ablock(plan, 4), niter=2)
    P64 = O.Problem(mumap, psfs, 0.8, angles32(nview).astype(np.float64), method, np.float64, "fft")
    xr = P64.osem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), 4, 2)
    assert np.linalg.norm(xg - xr) <= 1e-4 * np.linalg.norm(xr)
    plan.set_choice(psf_method=1)
    assert plan.info()["conv"] == "direct"
    gd = gpu.project(xtrue, plan)
    assert per_view_maxrel(gd, ytrue) <= 1e-5
    plan.set_choice(psf_method=0)
    rec = plan.autotune(2)
    cands = rec["psf_method"]["candidates"]
    assert [c["method"] for c in cands] == ["direct", "fft"] and all(c["fwd_us"] > 0 and c["adj_us"] > 0
                                                                     for c in cands)
    assert plan.choice()["psf_method"] == rec["psf_method"]["choice"]["code"]
    P32 = O.Problem(mumap, psfs, 0.8, angles32(nview).astype(np.float64), method, np.float32, "fft")
    assert per_view_maxrel(gpu.project(xtrue, plan), P32.project(xtrue)) <= PROJ_TOL
    zf = gpu.SPECTplan(mumap, np.asfortranarray(np.repeat(np.asarray(O.psf_gauss(nx, 5, 5, 1.0, 2.0), np.float32)
                                                          [:, :, :, None], nview, 3)), 0.8)
    if zf.info()["layout"] == "z-fastest":
        with pytest.raises(gpu.DeviceError):
            zf.set_choice(psf_method=2)


def _serve_oracle(O, cfg, req):
    """The FP64 oracle's OS-EM of a served / batched request (the plan's
    phantom mu-map and Gaussian PSFs, strided subsets as sr_osem_strided)."""
    import spectrecon_amd as S
    mumap = S.mumap_phantom(cfg.nx, cfg.nx, cfg.nz)
    psf = S.psf_gauss(ny=cfg.nx, px=cfg.psf, fwhm_start=cfg.fwhm_start, fwhm_end=cfg.fwhm_end)
    psfs = np.asfortranarray(np.repeat(psf[:, :, :, None], cfg.nview, 3))
    P64 = O.Problem(mumap, psfs, cfg.dy, S.default_viewangle(cfg.nview).astype(np.float64), "two", np.float64,
                    "fft", nthread=16)
    return P64.osem(req.x0.astype(np.float64), req.y.astype(np.float64), req.b.astype(np.float64), req.nblocks,
                    req.niter)


def test_batched_instances_vs_alone_and_oracle(gpu, O):
    """The batched uneven workload: the 32 seeded instances (nx 64/96/128, nz
    32/64/80, 60/120/128 views, PSF 15/25/37, 4/8 subsets) plus one at the
    clinical geometry (128x128x80, 128 views, 37x37, 8 subsets), each on its
    own resident plan with HIP graphs on, OS-EM x 4 from one sr_osem_batch
    call, against the same instance run alone (one call each, on its plan's
    stream): with the subset steps as shared launches (eager, graph capture,
    replay; the shared forward launches take the group's plane chunks, so
    the view sums round differently) within 1e-5 relative L2, the same on
    every run; with every instance on its own stream concurrently, bit for
    bit.  The clinical instance and an uneven 60-view / 8-subset instance
    match the FP64 oracle to 1e-4 in both forms."""
    from spectrecon_amd import serve as V
    inst = V.batched_instances() + [(V.Config(128, 80, 128, 37, 0.48, 1.0, 4.0), 8)]
    backend = V.DeviceBackend()
    plans, reqs = [], []
    for k, (cfg, nb) in enumerate(inst):
        plan = backend.make_plan(cfg)
        plans.append(plan)
        reqs.append(V.synthetic_request(cfg, seed=1 + k, nblocks=nb, niter=4, rid=k, plans={cfg: plan}))
    print(f"batched: {len(reqs)} instances prepared", flush=True)
    batch = V.Batch(plans, reqs)
    batch.run(sequential=True, streams=True)  # alone: each on its plan's stream (eager)
    alone = batch.results()
    s0 = V.batch_stats()
    first = None
    for rnd in range(3):  # shared launches: eager, graph capture, replay
        batch.run()
        res = batch.results()
        for k, (a, b) in enumerate(zip(alone, res)):
            rel = float(np.linalg.norm(a - b) / np.linalg.norm(a))
            assert rel <= 1e-5, (k, rnd, rel)
        if first is None:
            first = res
        for k, (a, b) in enumerate(zip(first, res)):
            np.testing.assert_array_equal(a, b, err_msg=f"shared launches, instance {k}, round {rnd}")
    shared_res = res
    s1 = V.batch_stats()
    shared = (s1["instances"] - s0["instances"]) // 3
    print(f"batched: {shared} of {len(reqs)} instances in shared launches, "
          f"{(s1['launches'] - s0['launches']) // 3} grouped launches per run", flush=True)
    assert shared >= len(reqs) // 2
    assert (s1["captures"] - s0["captures"], s1["replays"] - s0["replays"]) == (1, 1)
    for rnd in range(2):  # every instance on its own stream: plan graph capture, replay
        batch.run(streams=True)
        res = batch.results()
        for k, (a, b) in enumerate(zip(alone, res)):
            np.testing.assert_array_equal(a, b, err_msg=f"streams, instance {k}, round {rnd}")
    assert all(p.graph_stats() == (1, 1) for p in plans), [p.graph_stats() for p in plans]
    batch.free()
    uneven = next(k for k, (c, nb) in enumerate(inst) if c.nview == 60 and nb == 8)
    for k in (len(inst) - 1, uneven):
        xr = _serve_oracle(O, inst[k][0], reqs[k])
        for form, r in (("streams", res), ("shared", shared_res)):
            rel = float(np.linalg.norm(r[k] - xr) / np.linalg.norm(xr))
            print(f"batched instance {k} {inst[k]} ({form}): vs FP64 {rel:.2e}", flush=True)
            assert rel <= 1e-4, (k, form, rel)
    for p in plans:
        backend.close_plan(p)


def test_served_stream_vs_alone_and_oracle(gpu, O):
    """The served stream at 20 requests/s, Poisson arrivals, 70% clinical
    (128x128x80, 128 views, 37x37 PSF FWHM 1.5 -> 6) and 30% small (64x64x32,
    60 views, 15x15, FWHM 1 -> 4), OS-EM 8 x 4, through Server with the
    device backend (pinned staging, asynchronous copies and reconstructions,
    no host thread per request, graphs on): all 40 served, none rejected,
    every result equal to the same request reconstructed alone, and one
    request of each size within 1e-4 of the FP64 oracle."""
    from spectrecon_amd import serve as V
    big = V.Config(128, 80, 128, 37, 0.48, 1.5, 6.0)
    small = V.Config(64, 32, 60, 15, 0.48, 1.0, 4.0)
    rng = np.random.default_rng(11)
    plans, reqs, t = {}, [], 0.0
    for k in range(40):
        t += float(rng.exponential(1 / 20.0))
        cfg = big if rng.random() < 0.7 else small
        reqs.append(V.synthetic_request(cfg, seed=1000 + k, nblocks=8, niter=4, arrival=t, rid=k, plans=plans))
    for p in plans.values():
        p.close()
    server = V.Server(hbm_budget=100e9, max_active=8)
    try:
        st = server.run(reqs)
    finally:
        server.close()
    print(f"served: {st['served']} served, p50 {st['latency_ms']['p50']:.1f} ms, "
          f"p99 {st['latency_ms']['p99']:.1f} ms", flush=True)
    assert st["served"] == 40 and st["rejected"] == 0
    backend = V.DeviceBackend(graphs=False)
    alone_plans = {cfg: backend.make_plan(cfg) for cfg in (big, small)}
    for r in reqs:
        np.testing.assert_array_equal(r.result, gpu.osem_strided(r.x0, r.y, r.b, alone_plans[r.config], r.nblocks,
                                                                 r.niter), err_msg=f"request {r.rid}")
    for cfg in (big, small):
        r = next(q for q in reqs if q.config == cfg)
        xr = _serve_oracle(O, cfg, r)
        rel = float(np.linalg.norm(r.result - xr) / np.linalg.norm(xr))
        print(f"served request {r.rid} ({cfg.nx}x{cfg.nx}x{cfg.nz}): vs FP64 {rel:.2e}", flush=True)
        assert rel <= 1e-4, rel
    for p in alone_plans.values():
        p.close()
