"""GPU parity of the differentiable constant-kernel model and its adjoint.

Bars (tolerances are written per test):
  * chi^2 within 1e-5 relative, the gradient within 2e-4 of its largest entry
    and D within 2e-6 of the image scale, against the FP64 oracle, on ragged
    frames, every instantiated radius, with and without weights;
  * the GPU gradient equals central finite differences (step 1e-3) of the
    GPU chi^2 to 1e-3 of its largest entry (FP32 model arithmetic);
  * two independent GPU routes agree: the fused chi^2 adjoint and the
    vector-Jacobian product of D with G = 2 w D (also through autograd);
  * L-BFGS on the adjoint gradient recovers the generating kernel's beta and
    sub-pixel shift on a star field whose science PSF is the reference PSF
    convolved with a known shifted Gauss-Hermite kernel.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def nc_of(n_max):
    return (n_max + 1) * (n_max + 2) // 2


def star_pair(rng, h, w, n_stars=12):
    yy, xx = np.mgrid[0:h, 0:w]
    ref = np.full((h, w), 1000.0)
    sci = np.full((h, w), 1003.0)
    for _ in range(n_stars):
        x, y = rng.uniform(0, w), rng.uniform(0, h)
        f = 10 ** rng.uniform(3, 5)
        ref += f * np.exp(-0.5 * ((xx - x) ** 2 + (yy - y) ** 2) / 1.1 ** 2) / (2 * np.pi * 1.21)
        sci += f * np.exp(-0.5 * ((xx - x) ** 2 + (yy - y) ** 2) / 1.6 ** 2) / (2 * np.pi * 2.56)
    ref += 30 * rng.standard_normal((h, w))
    sci += 30 * rng.standard_normal((h, w))
    return sci.astype(np.float32), ref.astype(np.float32)


def params_at(rng, n_max, beta=1.35, dx=0.25, dy=-0.15, b=1.5):
    nc = nc_of(n_max)
    p = np.zeros(nc + 4)
    p[:nc] = 0.01 * rng.standard_normal(nc)
    p[0] = 0.21
    p[nc:] = (b, beta, dx, dy)
    return p


CASES = [  # (h, w, n_max, radius, weighted)
    (200, 137, 6, 15, True),
    (64, 64, 4, 11, True),
    (100, 90, 3, 7, False),
    (131, 70, 6, 12, True),
    (90, 75, 5, 9, False),
    (20, 17, 0, 1, True),
]


@pytest.mark.parametrize("h,w,n_max,r,weighted", CASES)
def test_chi2_gradient_difference_vs_fp64_oracle(gpu, h, w, n_max, r, weighted):
    import torch

    from delta_amd import _core

    rng = np.random.default_rng(h * 7 + w)
    sci, ref = star_pair(rng, h, w)
    wt = None
    if weighted:
        wt = (1.0 / (np.abs(sci) / 1.5 + np.abs(ref) / 1.5 * 0.05 + 7.1)).astype(np.float32)
        wt[rng.random((h, w)) < 0.05] = 0.0
    p = params_at(rng, n_max)
    print(f"case {h}x{w} n_max={n_max} r={r} weighted={weighted}", flush=True)
    got = _core.kernel_chi2(sci, ref, wt, p, n_max, r)
    o64 = oracle.kernel_chi2(sci, ref, wt, p, n_max, r, diff=True)
    assert abs(got["chi2"] / o64["chi2"] - 1) < 1e-5, (got["chi2"], o64["chi2"])
    err = np.abs(got["grad"] - o64["grad"]).max() / np.abs(o64["grad"]).max()
    assert err < 2e-4, (err, got["grad"], o64["grad"])
    d = _core.kernel_difference_device(torch.from_numpy(sci).to(gpu), torch.from_numpy(ref).to(gpu),
                                       torch.from_numpy(p).to(gpu), n_max, r).cpu().numpy()
    scale = np.abs(sci).max()
    assert np.abs(d - o64["difference"]).max() < 2e-6 * scale


def test_chi2_without_gradient_matches(gpu):
    from delta_amd import _core

    rng = np.random.default_rng(4)
    sci, ref = star_pair(rng, 150, 160)
    p = params_at(rng, 6)
    a = _core.kernel_chi2(sci, ref, None, p, 6, 15, grad=True)
    b = _core.kernel_chi2(sci, ref, None, p, 6, 15, grad=False)
    assert b["grad"] is None
    assert a["chi2"] == b["chi2"]


def test_gradient_matches_gpu_finite_differences(gpu):
    from delta_amd import _core

    rng = np.random.default_rng(21)
    h = w = 256
    sci, ref = star_pair(rng, h, w, n_stars=40)
    wt = (1.0 / (np.abs(sci) / 1.5 + 7.1)).astype(np.float32)
    n_max, r = 6, 15
    p = params_at(rng, n_max)
    g = _core.kernel_chi2(sci, ref, wt, p, n_max, r)["grad"]
    step = 1e-3
    fd = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = step
        fd[i] = (_core.kernel_chi2(sci, ref, wt, p + e, n_max, r, grad=False)["chi2"] -
                 _core.kernel_chi2(sci, ref, wt, p - e, n_max, r, grad=False)["chi2"]) / (2 * step)
    err = np.abs(g - fd) / np.abs(fd).max()
    assert err.max() < 1e-3, err


def test_vjp_and_autograd_agree_with_fused_adjoint(gpu):
    import torch

    from delta_amd import _core, adjoint

    rng = np.random.default_rng(8)
    h, w, n_max, r = 180, 200, 6, 15
    sci, ref = star_pair(rng, h, w, n_stars=20)
    wt = (1.0 / (np.abs(sci) / 1.5 + 7.1)).astype(np.float32)
    wt[:r] = 0
    S, R, W = (torch.from_numpy(a).to(gpu) for a in (sci, ref, wt))
    p = torch.from_numpy(params_at(rng, n_max)).to(gpu)
    fused = _core.kernel_chi2_device(S, R, W, p, n_max, r)
    d = _core.kernel_difference_device(S, R, p, n_max, r)
    vjp = _core.kernel_difference_vjp_device(R, (2 * W * d).contiguous(), p, n_max, r)
    g_f, g_v = fused[1:].cpu().numpy(), vjp[1:].cpu().numpy()
    assert np.abs(g_f - g_v).max() < 1e-5 * np.abs(g_f).max()
    # autograd through both Functions
    q = p.clone().requires_grad_(True)
    chi2 = adjoint.kernel_chi2(q, S, R, W, n_max, r)
    chi2.backward()
    g_a = q.grad.cpu().numpy()
    q2 = p.clone().requires_grad_(True)
    dd = adjoint.kernel_difference(q2, S, R, n_max, r)
    (W.double() * dd.double() ** 2).sum().backward()
    assert abs(float(chi2) / float(fused[0]) - 1) < 1e-12
    assert np.abs(g_a - g_f).max() <= 1e-12 * np.abs(g_f).max()
    assert np.abs(q2.grad.cpu().numpy() - g_f).max() < 1e-5 * np.abs(g_f).max()


def test_lbfgs_recovers_beta_and_shift(gpu):
    import torch

    from delta_amd import adjoint, synth

    n_max, r = 6, 15
    pair = synth.make_kernel_pair(size=512, seed=3, n_max=n_max, radius=r)
    truth = pair["params"]
    nc = nc_of(n_max)
    S, R = (torch.from_numpy(pair[k]).to(gpu) for k in ("science", "reference"))
    W = adjoint.chi2_weight(torch.from_numpy(pair["science_var"]).to(gpu),
                            torch.from_numpy(pair["reference_var"]).to(gpu), r,
                            kernel_power=0.05)
    p0 = np.zeros(nc + 4)
    p0[nc + 1] = 1.15
    p0[0] = 1.0
    p0[0] /= adjoint.shifted_kernel(p0, n_max, r).sum()
    free = np.zeros(nc + 4, bool)
    free[:15] = True          # orders <= 4: the generating kernel's order
    free[nc:] = True          # b, beta, dx, dy
    res = adjoint.refine_kernel(S, R, W, p0, n_max, r, free=free, max_iter=500)
    p = res["params"].cpu().numpy()
    print(f"beta {p[nc + 1]:.4f} dx {p[nc + 2]:.4f} dy {p[nc + 3]:.4f} "
          f"evaluations {res['evaluations']}", flush=True)
    assert abs(p[nc + 1] - truth[nc + 1]) < 0.02
    assert abs(p[nc + 2] - truth[nc + 2]) < 0.03
    assert abs(p[nc + 3] - truth[nc + 3]) < 0.03
    chi_truth = float(adjoint.kernel_chi2(torch.from_numpy(truth).to(gpu), S, R, W, n_max, r))
    assert res["chi2"] <= chi_truth * (1 + 1e-6)


def test_argument_errors(gpu):
    from delta_amd import _core

    sci = np.zeros((32, 32), np.float32)
    p = np.zeros(nc_of(6) + 4)
    p[nc_of(6) + 1] = 1.0
    with pytest.raises(ValueError):
        _core.kernel_chi2(sci, sci, None, p, 6, 16)
    with pytest.raises(ValueError):
        _core.kernel_chi2(sci, sci, None, np.zeros(nc_of(11) + 4), 11, 5)
    with pytest.raises(ValueError):
        _core.kernel_chi2(sci, sci, None, p[:-1], 6, 5)
    bad = p.copy()
    bad[nc_of(6) + 1] = 0.0
    with pytest.raises(ValueError):
        _core.kernel_chi2(sci, sci, None, bad, 6, 5)
