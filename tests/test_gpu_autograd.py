"""The soft pass as a torch.autograd.Function (nwgrad_amd.autograd): the adjoint equals the
batch's own weighted gradient, matches float64 central finite differences of the scores, and a
training step captured in a HIP graph replays to the same parameters as the eager steps."""
import numpy as np
import pytest

from nwgrad_amd import workloads as wl
from tests.gpu_util import from_workload, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _workload(name, n, related=0.5, lens=(30, 160)):
    return wl.random_pairs(name, n, wl.AA, lens, lens, related=related, grad_mode="soft")


def test_adjoint_is_weighted_grad(gpu):
    from nwgrad_amd import autograd as ag
    w = _workload("autograd_vjp", 300)
    b = from_workload(w, grad_mode="soft")
    theta = ag.params_vector(b_params(w)).requires_grad_(True)
    s = ag.soft_scores(theta, b)
    assert ag.finish(b) == 0
    wts = torch.tensor(wl.rng("autograd_vjp", "w").normal(size=w.n_pairs), dtype=torch.float64, device="cuda")
    (g,) = torch.autograd.grad((s * wts).sum(), theta)
    # the same batch through the host API: identical scores and the identical blocked sum
    np.testing.assert_array_equal(s.detach().cpu().numpy(), b.scores())
    want = b.weighted_grad(wts.cpu().numpy()).to_dict()
    flat = np.concatenate([np.asarray(want["matrix"]).ravel(), [want["gap_open_a"], want["gap_extend_a"],
                                                                 want["gap_open_b"], want["gap_extend_b"]]])
    np.testing.assert_array_equal(g.cpu().numpy(), flat)


def b_params(w):
    import nwgrad_amd as nw
    return nw.AlignParams(w.matrix, alphabet=w.alphabet, gap_open_a=w.gaps[0], gap_extend_a=w.gaps[1],
                          gap_open_b=w.gaps[2], gap_extend_b=w.gaps[3])


@pytest.mark.parametrize("mode", ["global", "local"])
def test_finite_differences_float64(gpu, mode):
    """Central differences (h = 1e-4) of sum_p w_p s_p on 16 random matrix cells and the 4 gap
    fields, float64 pass, against the adjoint."""
    from nwgrad_amd import autograd as ag
    w = wl.random_pairs(f"autograd_fd_{mode}", 64, wl.AA, (20, 90), (20, 90), related=0.5, mode=mode,
                        grad_mode="soft")
    b = from_workload(w, grad_mode="soft", precision="float64")
    theta0 = ag.params_vector(b_params(w))
    wts = torch.tensor(wl.rng(f"autograd_fd_{mode}", "w").normal(size=w.n_pairs), dtype=torch.float64,
                       device="cuda")

    def f(th):
        with torch.no_grad():
            v = float((ag.soft_scores(th, b) * wts).sum())
        assert ag.finish(b) == 0
        return v

    theta = theta0.clone().requires_grad_(True)
    (g,) = torch.autograd.grad((ag.soft_scores(theta, b) * wts).sum(), theta)
    g = g.cpu().numpy()
    nn = 400
    r = wl.rng(f"autograd_fd_{mode}", "cells")
    idx = np.concatenate([r.choice(nn, size=16, replace=False), nn + np.arange(4)])
    h = 1e-4
    worst = 0.0
    for k in idx:
        e = torch.zeros_like(theta0)
        e[k] = h
        fd = (f(theta0 + e) - f(theta0 - e)) / (2 * h)
        err = abs(fd - g[k]) / max(1.0, abs(g[k]))
        worst = max(worst, err)
        assert err <= 1e-6, (int(k), fd, g[k])
    say(f"  finite differences ({mode}): worst relative error {worst:.2e} over {len(idx)} fields")


def test_captured_step_replays(gpu):
    """A contrastive training step (soft pass, margin loss, adjoint, SGD update) captured in a
    graph: every replay, with eager work and synchronisation between replays, gives the
    parameters of the same number of eager steps."""
    from nwgrad_amd import autograd as ag
    w = wl.protein_soft_homologs(64, 64)
    b = from_workload(w, grad_mode="soft")
    n_pos, lr, margin = 64, 1e-3, 20.0

    def step(th):
        s = ag.soft_scores(th, b)
        loss = ag.margin_loss(s, n_pos, margin)
        (g,) = torch.autograd.grad(loss, th)
        with torch.no_grad():
            th.sub_(lr * g)
        return loss

    theta0 = ag.params_vector(b_params(w))
    eager = []
    th = theta0.clone().requires_grad_(True)
    for _ in range(4):
        step(th)
        torch.cuda.synchronize()
        eager.append(th.detach().clone())
    assert ag.finish(b) == 0

    th = theta0.clone().requires_grad_(True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step(th)   # warm-up: sizes the batch's device scratch outside the capture
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(th.detach().cpu().numpy(), eager[0].cpu().numpy())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step(th)
    for k in range(1, 4):
        graph.replay()
        torch.cuda.synchronize()
        _ = (torch.ones(8, device="cuda") * 2).sum().item()   # eager work between replays
        np.testing.assert_allclose(th.detach().cpu().numpy(), eager[k].cpu().numpy(), rtol=0, atol=1e-12)
    assert ag.finish(b) == 0


def test_training_reduces_loss(gpu):
    from nwgrad_amd import autograd as ag
    w = wl.protein_soft_homologs(96, 96)
    b = from_workload(w, grad_mode="soft")
    th = ag.params_vector(b_params(w)).requires_grad_(True)
    losses = []
    for _ in range(15):
        s = ag.soft_scores(th, b)
        loss = ag.margin_loss(s, 96, 40.0)
        (g,) = torch.autograd.grad(loss, th)
        with torch.no_grad():
            th.sub_(1e-2 * g)
        losses.append(float(loss))
    assert ag.finish(b) == 0
    say(f"  margin loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < losses[0]
