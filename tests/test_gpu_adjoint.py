"""GPU check of the score gradients against common-random-number central
finite differences (relative step 1e-3), on the two-media gradient workload
(Mie over Henyey-Greenstein): total reflectance and the summed coherent and
incoherent CBS S0, with respect to mu_s and mu_a of both layers and g of the
HG layer.  Sigma of both estimators from batch means over disjoint photon
ranges; the two must agree within a few combined sigma.

The gradients come from the production FP32 transport; the differences from
the FP64 transport on the same per-photon random streams: a relative step of 1e-3 moves
1/mu_t and mu_s/mu_t by a few FP32 ulps only, so FP32 rounding of the
perturbed coefficients would bias the difference quotient by several percent."""
import os

import numpy as np
import pytest

import luminis_mc as lmc
import scenes
from luminis_mc.adjoint import perturbed

pytestmark = pytest.mark.gpu

PARAMS = [(0, 0), (0, 1), (1, 0), (1, 1), (1, 2)]
NAMES = ["mu_s[0]", "mu_a[0]", "mu_s[1]", "mu_a[1]", "g[1]"]


def _quantities(v):
    """(R, sum S0 coherent, sum S0 incoherent) from a flat output vector."""
    return np.array([v[0], v[1:101].sum(), v[401:501].sum()])


def test_adjoint_matches_crn_finite_differences(require_gpu):
    sample, laser, sensors, keep = scenes.scene_adjoint()
    B, n = 12, 2_000_000
    tr = lmc.DifferentiableTransport(sample, laser, sensors, n_photons=n, seed=2024)
    tr64 = lmc.DifferentiableTransport(sample, laser, sensors, n_photons=n, seed=2024, precision="fp64")
    th0 = tr.theta0()
    s0 = tr.scene(th0)
    fd_scenes = [tuple(tr64.scene(t) for t in perturbed(th0, l, k)) for (l, k) in PARAMS]
    adj = np.zeros((B, len(PARAMS), 3))
    fd = np.zeros((B, len(PARAMS), 3))
    for b in range(B):
        _, jac = tr.evaluate(th0, with_grad=True, photon_offset=b * n, scene=s0)
        for j, (l, k) in enumerate(PARAMS):
            adj[b, j] = _quantities(jac[l, k])
            d = lmc.crn_central_difference(tr64, th0, l, k, photon_offset=b * n, scenes=fd_scenes[j])
            fd[b, j] = _quantities(d)
    ma, sa = adj.mean(0), adj.std(0, ddof=1) / np.sqrt(B)
    mf, sf = fd.mean(0), fd.std(0, ddof=1) / np.sqrt(B)
    z = (ma - mf) / np.sqrt(sa ** 2 + sf ** 2)
    for j, name in enumerate(NAMES):
        for q, qn in enumerate(["R", "S0coh", "S0inc"]):
            print(f"d{qn}/d{name}: adjoint {ma[j, q]:+.5e} +- {sa[j, q]:.1e}   FD {mf[j, q]:+.5e} +- {sf[j, q]:.1e}"
                  f"   z = {z[j, q]:+.2f}")
    print("fraction within 2 sigma:", np.mean(np.abs(z) < 2.0))
    assert np.all(np.abs(z) < 4.0)
    assert np.mean(np.abs(z) < 2.0) >= 0.7
    # the reflectance gradients are resolved: same sign, both well above their noise
    for j in range(len(PARAMS)):
        if abs(mf[j, 0]) > 5 * sf[j, 0]:
            assert np.sign(ma[j, 0]) == np.sign(mf[j, 0])
    s0.close()
    for a, b_ in fd_scenes:
        a.close()
        b_.close()


def test_autograd_backward_applies_the_jacobian(require_gpu):
    torch = pytest.importorskip("torch")
    sample, laser, sensors, keep = scenes.scene_adjoint()
    tr = lmc.DifferentiableTransport(sample, laser, sensors, n_photons=500_000, seed=7)
    th0 = tr.theta0()
    theta = torch.tensor(th0, requires_grad=True)
    out = tr(theta)
    assert out.dtype == torch.float64 and out.shape == (tr.n_out,)
    w = torch.zeros(tr.n_out, dtype=torch.float64)
    w[0] = 1.0
    w[1:101] = 0.5
    (out * w).sum().backward()
    vals, jac = tr.evaluate(th0, with_grad=True)
    expect = np.einsum("lpd,d->lp", jac, w.numpy())
    # two FP32 runs on the same photons agree to rounding: the order of the
    # FP64 atomic adds and, when a wave claims a second pool of photons at a
    # scheduling-dependent iteration, where the interior CBS product was last
    # renormalised (observed: <= 1e-4 relative in bins whose terms cancel)
    assert np.allclose(out.detach().numpy(), vals, rtol=1e-3, atol=1e-6 * np.abs(vals).max())
    assert np.allclose(theta.grad.numpy(), expect, rtol=1e-3, atol=1e-6 * np.abs(expect).max())
    # the Mie layer's g is not a parameter: its row is zero
    assert theta.grad[0, 2].item() == 0.0


def test_autograd_with_device_tensors_in_one_hip_runtime(require_gpu):
    """theta on the GPU: the library and torch share one HIP runtime (the
    loader binds the library to torch's), so the outputs, the Jacobian and the
    gradient stay on the device and equal the CPU-side evaluation."""
    torch = pytest.importorskip("torch")
    from luminis_mc import _lib
    assert torch.cuda.is_available()
    maps = open("/proc/self/maps").read()
    assert len({l.split()[-1] for l in maps.splitlines() if "libamdhip64" in l}) == 1
    if _lib.HIP_RUNTIME:
        assert os.path.realpath(_lib.HIP_RUNTIME) in maps
    sample, laser, sensors, keep = scenes.scene_adjoint()
    tr = lmc.DifferentiableTransport(sample, laser, sensors, n_photons=200_000, seed=8)
    th0 = tr.theta0()
    theta = torch.tensor(th0, device="cuda", requires_grad=True)
    out = tr(theta)
    assert out.is_cuda and out.dtype == torch.float64 and out.shape == (tr.n_out,)
    w = torch.zeros(tr.n_out, dtype=torch.float64, device="cuda")
    w[0] = 1.0
    w[1:101] = 0.25
    (out * w).sum().backward()
    assert theta.grad.is_cuda
    vals, jac = tr.evaluate(th0, with_grad=True)
    expect = np.einsum("lpd,d->lp", jac, w.cpu().numpy())
    assert np.allclose(out.detach().cpu().numpy(), vals, rtol=1e-3, atol=1e-6 * np.abs(vals).max())
    assert np.allclose(theta.grad.cpu().numpy(), expect, rtol=1e-3, atol=1e-6 * np.abs(expect).max())
