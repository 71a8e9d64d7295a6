"""The adjoint of the fit (mkf_fit_adjoint): forward chi2 against the oracle's FP32 model, the
gradient against FP64 central finite differences of the oracle's model, checkpoint independence,
the torch.autograd harness, and recovery of injected alignment offsets by gradient descent."""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from mkfit_amd.alignment import make_event, weak_mode_residual

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gpu_device):
    ev, geo, th_true = make_event(4000, seed=21)
    ctx = mk.Context(bfield=3.8)
    return ev, geo, th_true, ctx


def _theta(th, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(th, np.float32)).to(dev)


def test_fp64_forward_matches_fp64_oracle_model(setup, gpu_device):
    ev, geo, th, ctx = setup
    data = mk.AlignmentData(ctx, ev, gpu_device, fp64=True)
    chi2, _ = data.run(_theta(th, gpu_device))
    # the float32 theta the device sees, in the FP64 model
    c64 = oracle.align_chi2(ev, th.astype(np.float32).astype(np.float64), n_threads=8)
    g = chi2.cpu().numpy().astype(np.float64)
    rel = np.abs(g - c64) / np.maximum(1.0, c64)
    # per-track chi2 is returned as float32: agreement to its rounding
    assert np.percentile(rel, 99) < 1e-6 and rel.max() < 1e-4, (np.percentile(rel, 99), rel.max())


def test_forward_matches_oracle_model(setup, gpu_device):
    ev, geo, th, ctx = setup
    data = mk.AlignmentData(ctx, ev, gpu_device)
    chi2, _ = data.run(_theta(th, gpu_device))
    c32 = oracle.align_chi2(ev, th.astype(np.float32), dtype=np.float32, n_threads=8)
    c64 = oracle.align_chi2(ev, th, n_threads=8)
    g = chi2.cpu().numpy().astype(np.float64)
    # the GPU's FP32 model is as close to the FP64 model as the CPU's FP32 model is
    rel_g = np.abs(g - c64) / np.maximum(1.0, c64)
    rel_c = np.abs(c32 - c64) / np.maximum(1.0, c64)
    for q in (50, 90, 99):
        assert np.percentile(rel_g, q) <= 3 * np.percentile(rel_c, q) + 1e-6, (q, np.percentile(rel_g, q),
                                                                              np.percentile(rel_c, q))
    assert np.median(np.abs(g - c32) / np.maximum(1.0, c32)) < 1e-4
    assert abs(g.sum() - c64.sum()) / c64.sum() < 1e-4


@pytest.mark.parametrize("fp64", [True, False])
def test_reverse_sweep_matches_tangent_form(setup, gpu_device, fp64):
    """The step derivative by the hand-written reverse sweep against the same product from 30 forward
    tangents: in FP64 both are the exact derivative of the same model (agreement to rounding); in
    FP32 each is that derivative evaluated in float with its own operation order."""
    ev, geo, th, ctx = setup
    rev = mk.AlignmentData(ctx, ev, gpu_device, fp64=fp64)
    tan = mk.AlignmentData(ctx, ev, gpu_device, fp64=fp64, tangent=True)
    c_r, g_r = rev.run(_theta(th, gpu_device))
    c_t, g_t = tan.run(_theta(th, gpu_device))
    assert np.array_equal(c_r.cpu().numpy(), c_t.cpu().numpy())       # the same forward pass
    g_r, g_t = g_r.cpu().numpy().astype(np.float64), g_t.cpu().numpy().astype(np.float64)
    scale = np.abs(g_t).max(axis=0)                                   # per parameter kind
    d = np.abs(g_r - g_t) / scale
    assert d.max() < (1e-5 if fp64 else 2e-3), d.max(axis=0)


@pytest.mark.parametrize("fp64", [True, False])
def test_gradient_matches_fp64_finite_differences(setup, gpu_device, fp64):
    """FP64 adjoint: 1e-3 of the FP64 central differences (the required bound).  FP32 adjoint: the
    gradient of the FP32 model, whose residuals carry the rounding of two ~1 m coordinates (~6e-6 cm
    against ~1e-2 cm residuals); summed over tracks with cancellation that is a few 1e-3 relative."""
    ev, geo, th, ctx = setup
    th0 = th.copy()
    th0[:, 1:] += np.random.default_rng(3).normal(0, 20e-4, (th.shape[0], 2))   # off the optimum
    th0 = th0.astype(np.float32).astype(np.float64)      # the values the device sees
    data = mk.AlignmentData(ctx, ev, gpu_device, fp64=fp64)
    _, grad = data.run(_theta(th0, gpu_device))
    grad = grad.cpu().numpy().astype(np.float64)
    L = th.shape[0]
    h = 1e-6
    fd = np.zeros_like(grad)
    for l in range(L):
        for k in range(3):
            tp = th0.copy(); tp[l, k] += h
            tm = th0.copy(); tm[l, k] -= h
            fd[l, k] = (oracle.align_chi2(ev, tp, n_threads=8).sum() - oracle.align_chi2(ev, tm, n_threads=8).sum()) / (2 * h)
    bound = 1e-3 if fp64 else 1e-2
    for k in range(3):
        scale = np.abs(fd[:, k]).max()
        err = np.abs(grad[:, k] - fd[:, k]) / np.maximum(np.abs(fd[:, k]), 1e-2 * scale)
        print("fp64=%s param %d: max rel err %.3g (layer %d), median %.3g" % (fp64, k, err.max(), np.argmax(err),
                                                                              np.median(err)))
        assert err.max() < bound, (k, err.max(), np.argmax(err))


@pytest.mark.parametrize("fp64", [False, True])
def test_checkpointing_does_not_change_the_result(setup, gpu_device, fp64):
    ev, geo, th, ctx = setup
    out = []
    for ck in (1, 5, 25):
        data = mk.AlignmentData(ctx, ev, gpu_device, checkpoint_every=ck, fp64=fp64)
        c, g = data.run(_theta(th, gpu_device))
        out.append((c.cpu().numpy(), g.cpu().numpy()))
    for c, g in out[1:]:
        assert np.array_equal(c, out[0][0]) and np.array_equal(g, out[0][1])


def test_autograd_harness(setup, gpu_device):
    import torch
    from mkfit_amd.autograd import fit_chi2
    ev, geo, th, ctx = setup
    data = mk.AlignmentData(ctx, ev, gpu_device)
    theta = _theta(th, gpu_device).requires_grad_(True)
    loss = fit_chi2(theta, data)
    loss.backward()
    chi2, grad = data.run(theta.detach())
    assert torch.allclose(theta.grad, grad)
    assert abs(loss.item() - chi2.double().sum().item()) <= 1e-6 * abs(loss.item())


def test_gradient_descent_recovers_offsets(gpu_device):
    """20 diagonally preconditioned descent steps on the offsets, material held at its true value."""
    import torch
    ev, geo, th_true = make_event(50000, seed=22)
    ctx = mk.Context(bfield=3.8)
    data = mk.AlignmentData(ctx, ev, gpu_device)
    L = th_true.shape[0]
    radii = np.array([geo.barrel_r[l] for l in range(L)])
    th = th_true.copy()
    th[:, 1:] = 0.0
    # diagonal curvature from differences of the adjoint gradient
    H = np.zeros((L, 3))
    d = 1e-4
    for l in range(L):
        for k in (1, 2):
            tp = th.copy(); tp[l, k] += d
            tm = th.copy(); tm[l, k] -= d
            gp = data.run(_theta(tp, gpu_device))[1].cpu().numpy()
            gm = data.run(_theta(tm, gpu_device))[1].cpu().numpy()
            H[l, k] = (gp[l, k] - gm[l, k]) / (2 * d)
    assert np.all(H[:, 1:] > 0)
    for _ in range(20):
        g = data.run(_theta(th, gpu_device))[1].cpu().numpy().astype(np.float64)
        th[:, 1:] -= 0.5 * g[:, 1:] / H[:, 1:]
    r_rphi = weak_mode_residual(th[:, 1], th_true[:, 1], radii, "rphi")
    r_z = weak_mode_residual(th[:, 2], th_true[:, 2], radii, "z")
    s_rphi = weak_mode_residual(np.zeros(L), th_true[:, 1], radii, "rphi")
    s_z = weak_mode_residual(np.zeros(L), th_true[:, 2], radii, "z")
    assert r_rphi.std() < 0.1 * s_rphi.std(), (r_rphi.std(), s_rphi.std())
    assert r_z.std() < 0.3 * s_z.std(), (r_z.std(), s_z.std())


def test_two_streams_share_the_scratch_in_order(setup, gpu_device):
    """Two gradient calls of one context on two streams, queued back to back without a host wait: the
    second call's stream waits for the first call's kernels before it reuses the context's checkpoint
    and gradient-row scratch, so both results equal single-stream runs."""
    import torch
    ev, geo, th, ctx = setup
    a = mk.AlignmentData(ctx, ev, gpu_device)
    half = {k: np.ascontiguousarray(v[..., :2000]) for k, v in ev.items()}
    b = mk.AlignmentData(ctx, half, gpu_device)
    th_a, th_b = _theta(th, gpu_device), _theta(th * 1.01, gpu_device)
    ref_a = [x.cpu().numpy() for x in a.run(th_a)]
    ref_b = [x.cpu().numpy() for x in b.run(th_b)]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        with torch.cuda.stream(s1):        # run() launches and copies its results on the current stream
            out_a = a.run(th_a)
        with torch.cuda.stream(s2):
            out_b = b.run(th_b)
        torch.cuda.synchronize()
        assert np.array_equal(out_a[0].cpu().numpy(), ref_a[0]) and np.array_equal(out_a[1].cpu().numpy(), ref_a[1])
        assert np.array_equal(out_b[0].cpu().numpy(), ref_b[0]) and np.array_equal(out_b[1].cpu().numpy(), ref_b[1])
