"""CPU tests of the product library's boundary and host logic: the C-ABI
loads and exports every declared symbol, argument checks raise the
reference's messages before any device work, and the host-side generators
match the oracle and the reference's known answers."""
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spectrecon_amd.h")


def declared_symbols():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(sr_\w+)\(", src, flags=re.M)))


def test_library_exports_every_declared_symbol(S):
    from spectrecon_amd import _lib

    names = declared_symbols()
    assert len(names) >= 20
    L = S.lib()
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, n
    assert set(_lib.SIGNATURES) == set(names)


def test_header_constants_match_binding():
    from spectrecon_amd import _lib

    src = open(HEADER).read()
    consts = dict(re.findall(r"#define (SR_\w+) (0x[0-9a-f]+|\d+)", src))
    for name, val in consts.items():
        if hasattr(_lib, name):
            assert getattr(_lib, name) == int(val, 0), name


def test_version_and_device_count(S):
    assert S.lib().sr_version() > 0
    assert S.device_count() >= 0  # never fails, also without a GPU


def _plan_inputs(nx=8, ny=8, nz=4, px=3, pz=3, nview=4):
    mumap = np.zeros((nx, ny, nz), np.float32)
    psfs = np.ones((px, pz, ny, nview), np.float32) / (px * pz)
    return mumap, psfs


@pytest.mark.parametrize("kw,msg", [
    (dict(nx=8, ny=6), "nx != ny"),
    (dict(nx=7, ny=7), "nx odd"),
    (dict(px=4), "non-odd size psfs"),
    (dict(pz=2), "non-odd size psfs"),
])
def test_plan_checks_messages(S, kw, msg):
    """src/spectplan.jl:76-82: the same strings the reference throws, raised
    before any device work (so also without a GPU)."""
    mumap, psfs = _plan_inputs(**kw)
    with pytest.raises(S.SPECTError) as e:
        S.SPECTplan(mumap, psfs, 0.48)
    assert str(e.value) == msg


def test_plan_asymmetric_psf(S):
    mumap, psfs = _plan_inputs()
    psfs[0, 0, 2, 1] += 0.01
    with pytest.raises(S.SPECTError, match="asym. psf"):
        S.SPECTplan(mumap, psfs, 0.48)


@pytest.mark.parametrize("kw,msg", [(dict(interpmeth="three"), "bad interpmeth"), (dict(mode="slow"), "bad mode")])
def test_plan_option_messages(S, kw, msg):
    mumap, psfs = _plan_inputs()
    with pytest.raises(S.SPECTError, match=msg):
        S.SPECTplan(mumap, psfs, 0.48, **kw)


def test_plan_converts_inputs_in_place(S):
    """SPECTplan mutates mumap/psfs to T (src/spectplan.jl:71-72)."""
    mumap, psfs = _plan_inputs()
    mumap = mumap.astype(np.float64) + 0.1 + 1e-12
    try:
        S.SPECTplan(mumap, psfs, 0.48)
    except S.DeviceError:
        pass  # no GPU here: the conversion happened before device work
    assert np.array_equal(mumap, mumap.astype(np.float32).astype(np.float64))


def test_convenience_dimension_checks(S):
    """src/project.jl:204-205, src/backproject.jl:198-200."""
    mumap, psfs = _plan_inputs()
    with pytest.raises(S.DimensionMismatch, match="image/mumap size"):
        S.project(np.zeros((8, 8, 5), np.float32), mumap, psfs, 0.48)
    with pytest.raises(S.DimensionMismatch, match="image/psfs size"):
        S.project(np.zeros((8, 8, 4), np.float32), mumap, psfs[:, :, :6], 0.48)
    with pytest.raises(S.DimensionMismatch, match="nz"):
        S.backproject(np.zeros((8, 5, 4), np.float32), mumap, psfs, 0.48)


def test_fft_conv_asymmetric_kernel(S):
    """test/fftconv.jl:28: an asymmetric kernel throws."""
    rng = np.random.default_rng(0)
    with pytest.raises(S.SPECTError, match="asymmetric kernel"):
        S.fft_conv(rng.standard_normal((12, 8)), rng.random((7, 7)))


def test_plan_rotate_and_psf_pads(S, O):
    """padsize (src/plan-rotate.jl:56) and the PSF pads (src/helper.jl:8-12)."""
    for nx in (8, 16, 64, 96, 128, 140):
        assert S.plan_rotate(nx)[0].padsize == O.padsize(nx)
    assert S.plan_rotate(128)[0].padsize == 28 and S.plan_rotate(64)[0].padsize == 15
    p = S.plan_psf(nx=128, nz=80, px=37, pz=37)[0]
    assert p.padsize == (64, 64, 24, 24)
    assert p.padsize[:2] == O.psf_pads(128, 37) and p.padsize[2:] == O.psf_pads(80, 37)


def test_psf_gauss_product_matches_oracle(S, O):
    for ny, px, pz, a, b in [(4, 5, 3, 1.0, 4.0), (128, 37, 37, 1.5, 6.0), (64, 15, 15, 1.0, 4.0)]:
        assert np.array_equal(S.psf_gauss(ny=ny, px=px, pz=pz, fwhm_start=a, fwhm_end=b),
                              O.psf_gauss(ny, px, pz, a, b))
    ny, px, pz = 4, 5, 3
    p = S.psf_gauss(ny=ny, px=px, pz=pz, fwhm_x=np.full(ny, np.inf), fwhm_z=np.zeros(ny))
    tmp = np.zeros((px, pz))
    tmp[:, (pz + 1) // 2 - 1] = 1 / px
    assert np.allclose(p, np.repeat(tmp[:, :, None], ny, 2))
    assert S.psf_gauss().dtype == np.float32 and S.psf_gauss().shape == (11, 11, 128)


def test_phantom_generators(S):
    a = S.phantom(32, 32, 16, seed=3)
    assert np.array_equal(a, S.phantom(32, 32, 16, seed=3))
    assert not np.array_equal(a, S.phantom(32, 32, 16, seed=4))
    assert a.min() == 0 and 0.5 <= a.max() <= 6 * 4
    mu = S.mumap_phantom(64, 64, 32)
    assert set(np.unique(mu).tolist()) == {0.0, np.float32(0.05), np.float32(0.15)}
    # mid-plane: the inner ellipsoids sit at +-0.18 nx from the centre
    c = (64 + 1) / 2
    assert mu[int(round(c - 0.18 * 64)) - 1, 31, 15] == np.float32(0.05)
    assert mu[31, 31, 15] == np.float32(0.15)


def test_poisson_generator(S):
    lam = np.full(200000, 7.5, np.float32)
    s = S.poisson(lam, seed=1)
    assert np.array_equal(s, S.poisson(lam, seed=1))
    assert np.all(s == np.round(s)) and s.min() >= 0
    assert abs(s.mean() - 7.5) < 0.03 and abs(s.var() - 7.5) < 0.15
    lam = np.full(200000, 250.0, np.float32)  # PTRS branch
    s = S.poisson(lam, seed=2)
    assert abs(s.mean() - 250) < 0.3 and abs(s.var() / 250 - 1) < 0.03
    assert S.poisson(np.zeros(10, np.float32)).sum() == 0


def test_default_viewangle(S):
    va = S.default_viewangle(128)
    assert va.dtype == np.float32 and va[0] == 0
    assert np.all(np.diff(va) > 0) and va[-1] < np.float32(2 * math.pi)


def test_no_cpu_fallback(S, monkeypatch):
    """The product path has no CPU fallback: without the library it raises."""
    from spectrecon_amd import _lib

    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", os.path.join(ROOT, "does-not-exist.so"))
    with pytest.raises(ImportError):
        _lib.lib()


def _fake_plan(nx=8, nz=4, nview=4):
    """A stand-in with a plan's sizes and no device handle: the shape checks
    must raise before any library call would use it."""
    from types import SimpleNamespace

    return SimpleNamespace(imgsize=(nx, nx, nz), nview=nview, handle=None)


@pytest.mark.parametrize("which", ["x0", "out", "y", "b"])
def test_em_device_paths_check_shapes(S, which):
    """mlem_/osem_/osem_strided raise DimensionMismatch for arrays whose
    sizes differ from the plan's, before the library reads or writes them."""
    plan = _fake_plan()
    good = {"x0": np.ones((8, 8, 4), np.float32), "out": np.zeros((8, 8, 4), np.float32),
            "y": np.ones((8, 4, 4), np.float32), "b": np.ones((8, 4, 4), np.float32)}
    bad = dict(good)
    bad[which] = np.ones((8, 8, 3), np.float32) if which in ("x0", "out") else np.ones((8, 4, 3), np.float32)
    with pytest.raises(S.DimensionMismatch):
        S.mlem_(bad["out"], bad["x0"], bad["y"], bad["b"] if which != "y" else np.ones((8, 4, 3), np.float32),
                S.spect_operator(plan), niter=1)
    with pytest.raises(S.DimensionMismatch):
        S.osem_(bad["out"], bad["x0"], bad["y"], bad["b"], S.Ablock(plan, 2), niter=1)
    if which != "out":
        with pytest.raises(S.DimensionMismatch):
            S.osem_strided(bad["x0"], bad["y"], bad["b"], plan, 2, 1)
