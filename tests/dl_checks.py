"""Checks of the projector inside torch autograd and of the regularised EM
step of the reference's deep-learning example (docs/lit/examples/6-dl.jl:
150-198), run as a script in a process of its own: torch is imported before
the library, so that both use torch's HIP runtime (tests/test_gpu_dl.py runs
it; the pytest process has loaded the library already).

- the autograd passes are the library's own operators, bit for bit;
- bregem matches a restatement over the FP64 oracle;
- gradients reach a 3-D convolution prior and beta (d loss / d beta against
  a central difference).
Prints one "ok <check>" line per check; any failure raises."""
import os
import sys

import torch  # first: the library then shares torch's HIP runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "spectrecon.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
import spectrecon_amd as S  # noqa: E402
from spectrecon_amd import dl  # noqa: E402


def to_t(a):
    """numpy [nx, ny, nz] (column-major) -> tensor (nz, ny, nx) on the GPU."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).T)).cuda()


def from_t(t):
    return np.asfortranarray(t.detach().cpu().numpy().T)


def angles32(nview):
    return S.default_viewangle(nview)


def problem(nx=32, nz=16, nview=12):
    """The EM test problem of the parity tests: two boxes, a random mu-map, Gaussian PSFs."""
    rng = np.random.default_rng(2)
    x = np.zeros((nx, nx, nz), np.float32)
    x[nx // 4:2 * nx // 3, nx // 5:3 * nx // 5, 2 * nz // 6:3 * nz // 6] = 1
    x[2 * nx // 5:3 * nx // 5, nx // 5:2 * nx // 5, 4 * nz // 6:5 * nz // 6] = 2
    psf = O.psf_gauss(nx, 5, 5, 1.0, 4.0)
    psfs = np.asfortranarray(np.repeat(psf[:, :, :, None], nview, 3))
    mumap = np.asfortranarray((0.02 * rng.random((nx, nx, nz))).astype(np.float32))
    plan = S.SPECTplan(mumap, psfs, 0.8)
    ytrue = S.project(x, plan)
    b = np.full_like(ytrue, 0.1 * ytrue.mean())
    y = S.poisson(20 * (ytrue + b), seed=7) / 20
    return dict(xtrue=np.asfortranarray(x), mumap=mumap, psfs=psfs, plan=plan, y=y, b=b, nview=nview)


def check_autograd(p):
    plan, xtrue = p["plan"], p["xtrue"]
    g = torch.Generator(device="cuda").manual_seed(3)
    x = to_t(xtrue).requires_grad_()
    y = dl.project(x, plan)
    assert tuple(y.shape) == dl.views_shape(plan)
    np.testing.assert_array_equal(from_t(y), S.project(xtrue, plan))
    w = torch.rand(y.shape, generator=g, device="cuda")
    (w * y).sum().backward()
    np.testing.assert_array_equal(from_t(x.grad), S.backproject(from_t(w), plan))
    v = torch.rand(dl.views_shape(plan), generator=g, device="cuda").requires_grad_()
    bp = dl.backproject(v, plan)
    np.testing.assert_array_equal(from_t(bp), S.backproject(from_t(v), plan))
    m = torch.rand(bp.shape, generator=g, device="cuda")
    (m * bp).sum().backward()
    np.testing.assert_array_equal(from_t(v.grad), S.project(from_t(m), plan))
    for bad, exc in ((torch.zeros((3, 4, 5), device="cuda"), S.DimensionMismatch),
                     (torch.zeros(dl.image_shape(plan)), TypeError)):
        try:
            dl.project(bad, plan)
        except exc:
            pass
        else:
            raise AssertionError(f"no {exc.__name__}")


def check_bregem_vs_oracle(p):
    """Two inner EM iterations with a fixed prior u = 0.05 x^2 against the same
    closed-form update over the FP64 oracle's projector, relative L2 <= 1e-4."""
    plan, y, b, nview, beta = p["plan"], p["y"], p["b"], p["nview"], 0.5
    P64 = O.Problem(p["mumap"], p["psfs"], 0.8, angles32(nview).astype(np.float64), "two", np.float64, "direct")
    x0 = np.ones_like(p["xtrue"])
    ones = np.ones_like(y)
    xt = dl.bregem(to_t(y), to_t(b), dl.backproject(to_t(ones), plan), to_t(x0), lambda t: 0.05 * t * t, beta,
                   plan, niter=2)
    y64, b64, x = y.astype(np.float64), b.astype(np.float64), x0.astype(np.float64)
    asumu = P64.backproject(ones.astype(np.float64)) - beta * 0.05 * x * x
    for _ in range(2):
        eterm = P64.backproject(y64 / (P64.project(x) + b64))
        x = np.maximum(0, (-asumu + np.sqrt(asumu ** 2 + 4 * beta * x * eterm)) / (2 * beta))
    err = np.linalg.norm(from_t(xt) - x) / np.linalg.norm(x)
    assert err <= 1e-4, err


def check_training(p):
    plan, y, b = p["plan"], p["y"], p["b"]
    torch.manual_seed(0)
    conv = torch.nn.Conv3d(1, 1, 3, padding=1).cuda()

    def cnn(t):
        return conv(t[None, None])[0, 0]

    yt, rt, xtrue = to_t(y), to_t(b), to_t(p["xtrue"])
    asum = dl.backproject(torch.ones_like(yt), plan)
    x0 = torch.ones_like(xtrue)

    def loss_at(beta):
        xr = dl.bregem(yt, rt, asum, x0, cnn, beta, plan, niter=2)
        return torch.mean((xr - xtrue) ** 2)

    beta = torch.tensor(0.5, device="cuda", requires_grad=True)
    loss_at(beta).backward()
    grads = [q.grad for q in conv.parameters()] + [beta.grad]
    assert all(q is not None and bool(torch.isfinite(q).all()) for q in grads)
    assert any(float(q.abs().max()) > 0 for q in grads)
    h = 1e-2
    with torch.no_grad():
        fd = (loss_at(torch.tensor(0.5 + h, device="cuda")) - loss_at(torch.tensor(0.5 - h, device="cuda"))) / (2 * h)
    assert abs(float(beta.grad) - float(fd)) <= 2e-2 * abs(float(fd)) + 1e-7, (float(beta.grad), float(fd))


def main():
    assert torch.cuda.is_available(), "torch sees no GPU"
    p = problem()
    for name, fn in (("autograd", check_autograd), ("bregem_vs_oracle", check_bregem_vs_oracle),
                     ("training", check_training)):
        print(f"run {name}", flush=True)
        fn(p)
        print(f"ok {name}", flush=True)


if __name__ == "__main__":
    main()
