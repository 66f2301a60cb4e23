"""The learned preconditioner's network (MFMA BF16 convolutions) against a
PyTorch-CPU autograd oracle of the same network: forward output, loss and the
gradient of every parameter, then a training run.  The oracle rounds the
input, the weights and the stored hidden activations to BF16 exactly where
the kernels do and computes in FP64; the kernels also store the backward
signals in BF16, which the oracle does not, hence the gradient tolerance."""
import numpy as np
import pytest

sa = pytest.importorskip("seismic_amd")
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

C, TAPS, WL, PL = 16, 9, 9 * 16 * 16, 9 * 16 * 16 + 16


def bf16(a):
    """Round FP32 to BF16 (nearest even), returned as FP32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def torch_net(theta, x, round_acts=True):
    """x: (nimg, nz, nx) float64 tensor; theta: FP32 parameter tensor (requires grad)."""
    h = x[:, None]
    for l in range(4):
        W = theta[l * PL:l * PL + WL].reshape(TAPS, C, C)           # [t][co][ci]
        b = theta[l * PL + WL:(l + 1) * PL]
        Wq = W + (torch.from_numpy(bf16(W.detach().numpy())) - W).detach()   # BF16 value, identity gradient
        w = Wq.reshape(3, 3, C, C).permute(2, 3, 1, 0).double()       # [co][ci][dz+1][dx+1]
        ci = 1 if l == 0 else C
        co = 1 if l == 3 else C
        h = torch.nn.functional.conv2d(h, w[:co, :ci], b[:co].double(), padding=1)
        if l < 3:
            h = torch.relu(h)
            if round_acts:
                h = h + (torch.from_numpy(bf16(h.detach().float().numpy())).double() - h).detach()
    return h[:, 0]


def case(seed=0, nimg=2, nz=37, nx=53):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nimg, nz, nx)).astype(np.float32)
    t = (0.5 * x + 0.25 * np.roll(x, 1, axis=1)).astype(np.float32)
    return x, t


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_cnn_forward_and_gradient_match_torch():
    print("case: cnn forward/gradient vs torch", flush=True)
    x, t = case()
    net = sa.CNN(x.shape[1], x.shape[2], nimg_max=2, seed=3)
    th = net.params()
    y = net.forward(x)
    loss, g = net.loss_grad(x, t)
    th_t = torch.from_numpy(th.copy()).requires_grad_(True)
    xt = torch.from_numpy(bf16(x)).double()
    yt = torch_net(th_t, xt)
    lt = ((yt - torch.from_numpy(t).double()) ** 2).mean()
    lt.backward()
    gt = th_t.grad.numpy()
    ey, eg = rel(y, yt.detach().numpy()), rel(g, gt)
    print("forward rel %.2e  loss %.6e vs %.6e  grad rel %.2e" % (ey, loss, lt.item(), eg), flush=True)
    assert ey < 1e-2
    assert abs(loss - lt.item()) < 2e-2 * lt.item()
    assert eg < 3e-2
    for l in range(4):   # per layer, weights and biases separately
        for sl in (slice(l * PL, l * PL + WL), slice(l * PL + WL, (l + 1) * PL)):
            if np.linalg.norm(gt[sl]) > 0:
                assert rel(g[sl], gt[sl]) < 5e-2, (l, sl)
    # padding of the 1-channel layers carries no gradient into the parameters
    th2 = th.copy()
    th2[0:WL].reshape(TAPS, C, C)[:, :, 1:] = 1.0
    net.set_params(th2)
    assert np.array_equal(net.params(), th)


def test_cnn_training_reduces_loss_and_matches_torch_loss():
    print("case: cnn training", flush=True)
    x, t = case(seed=1)
    net = sa.CNN(x.shape[1], x.shape[2], nimg_max=2, seed=5)
    l0, l1 = net.train(x, t, steps=200, lr=3e-3)
    print("loss %.4e -> %.4e" % (l0, l1), flush=True)
    assert l1 < 0.1 * l0
    th = torch.from_numpy(net.params())
    yt = torch_net(th, torch.from_numpy(bf16(x)).double())
    lt = ((yt - torch.from_numpy(t).double()) ** 2).mean().item()
    assert abs(lt - l1) < 3e-2 * max(lt, 1e-12) + 1e-6


def _fwi_case():
    from models import ricker, smooth_random
    nz, nx, nt = 40, 96, 300
    true = smooth_random(nz, nx, seed=21, vmin=1600, vmax=2600)
    init = smooth_random(nz, nx, seed=22, vmin=1600, vmax=2500)
    w = ricker(18.0, nt, 1e-3)
    shots = []
    for sx in [200.0, 470.0, 760.0]:
        rx = np.linspace(max(0.0, sx - 300), min(10.0 * (nx - 1), sx + 300), 41).astype(np.float32)
        shots.append(sa.Shot(12.0, sx, w, np.full_like(rx, 15.0), rx))
    e = sa.Engine(sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12))
    e.set_model(true)
    dobs = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    return e, shots, dobs, init


def test_fwi_with_learned_preconditioner_runs_gated():
    """FWI with the CNN preconditioner: the network is trained on the GPU from
    the first iterations' pairs, then used as H0; the misfit still decreases,
    and the reported gain over plain L-BFGS (same evaluations budget) is printed."""
    print("case: fwi with learned preconditioner", flush=True)
    e, shots, dobs, init = _fwi_case()
    kw = dict(niter=10, neval=40, bound=0, illupow=0.0)
    _, rp = e.fwi_lbfgs(shots, dobs, init, **kw)
    _, rc = e.fwi_lbfgs(shots, dobs, init, precond=dict(train_after=3, train_steps=100), **kw)
    fp = rp["f_eval"][rp["accepted"]][-1] / rp["f_eval"][0]
    fc = rc["f_eval"][rc["accepted"]][-1] / rc["f_eval"][0]
    print("plain: %d iterations, final/initial misfit %.4e; preconditioned: %d iterations (%d with the network, "
          "%d restarts), %.4e; training loss %.3e -> %.3e" % (rp["iterations"], fp, rc["iterations"],
                                                             rc["pc_iterations"], rc["pc_fallbacks"], fc,
                                                             *rc["pc_loss"]), flush=True)
    assert rc["pc_train_seconds"] > 0 and rc["pc_loss"][1] < rc["pc_loss"][0]
    assert rc["iterations"] >= 4 and fc < 1.0
    assert np.all(np.diff(rc["f_eval"][rc["accepted"]]) < 0)


def test_preconditioned_search_falls_back_to_plain():
    """A preconditioned line search allowed no evaluation is abandoned at once
    and restarted plain from the same model (restart_plain); after
    max_fallbacks restarts the network is switched off.  Every direction then
    is the plain one, so the run must retrace plain L-BFGS exactly."""
    print("case: preconditioner fallback", flush=True)
    e, shots, dobs, init = _fwi_case()
    kw = dict(niter=8, neval=40, bound=0, illupow=0.0)
    xp, rp = e.fwi_lbfgs(shots, dobs, init, **kw)
    xc, rc = e.fwi_lbfgs(shots, dobs, init, precond=dict(train_after=2, train_steps=50, max_evals=0,
                                                         max_fallbacks=2), **kw)
    print("iterations %d/%d, restarts %d, with network %d" % (rc["iterations"], rp["iterations"],
                                                              rc["pc_fallbacks"], rc["pc_iterations"]), flush=True)
    assert rc["pc_fallbacks"] == 2 and rc["pc_iterations"] == 0
    assert rc["evaluations"] == rp["evaluations"] and rc["iterations"] == rp["iterations"]
    assert np.array_equal(rc["f_eval"], rp["f_eval"])
    assert np.array_equal(xc, xp)


@pytest.mark.parametrize("use_net", [1, 0])
def test_coarse_grid_preconditioner(use_net):
    """The network on the 4x coarser grid (H0 = g P N(R q), trained on R g_k ->
    R s_k), or with use_net 0 the coarse-grid smoother P R alone, serving from
    the first iteration: the run descends at every accepted step and the
    operator is used."""
    print("case: coarse-grid preconditioner use_net=%d" % use_net, flush=True)
    e, shots, dobs, init = _fwi_case()
    kw = dict(niter=8, neval=32, bound=0, illupow=0.0)
    _, rc = e.fwi_lbfgs(shots, dobs, init, precond=dict(train_after=3, train_steps=100, coarsen=4,
                                                         use_net=use_net), **kw)
    fc = rc["f_eval"][rc["accepted"]]
    print("iterations %d, with the operator %d, restarts %d, final/initial %.4e, loss %s" % (
        rc["iterations"], rc["pc_iterations"], rc["pc_fallbacks"], fc[-1] / rc["f_eval"][0], rc["pc_loss"]),
        flush=True)
    assert rc["iterations"] >= 4 and fc[-1] < 0.5 * rc["f_eval"][0]
    assert np.all(np.diff(fc) < 0)
    assert rc["pc_iterations"] >= (2 if use_net else 4)
    if use_net:
        assert rc["pc_train_seconds"] > 0 and rc["pc_loss"][1] < rc["pc_loss"][0]
