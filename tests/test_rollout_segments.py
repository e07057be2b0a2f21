"""CPU: the graph-free checkpointed segments of ``checkpointed_rollout``
(``params`` given) against the same rollout without checkpoints, on a toy
step that reads a parameter from its closure: equal values, equal gradients
with respect to the initial state and the parameter, and no graph kept
inside a segment between the forward and the backward pass."""

from __future__ import annotations

import torch

from gridfoam_amd.adjoint import checkpointed_rollout


def _problem():
    gen = torch.Generator().manual_seed(3)
    x0 = torch.randn(50, dtype=torch.float64, generator=gen)
    y0 = torch.randn(50, dtype=torch.float64, generator=gen)
    a = torch.tensor(0.7, dtype=torch.float64, requires_grad=True)
    b = torch.randn(50, dtype=torch.float64, generator=gen).requires_grad_()
    P = {}
    calls = []

    def fresh_w():
        P["w"] = b * 2.0  # a non-leaf parameter: the rollout stops at it, autograd carries on to b
        return P["w"]

    def step(s):
        x, y = s
        calls.append(torch.is_grad_enabled())
        return torch.sin(x) * a + y * 0.5, x * P["w"] - 0.1 * y ** 2

    return x0, y0, a, b, fresh_w, step, calls


def test_segments_match_the_full_graph():
    x0, y0, a, b, fresh_w, step, calls = _problem()
    xr, yr = x0.clone().requires_grad_(), y0.clone().requires_grad_()
    fresh_w()
    out = checkpointed_rollout(step, (xr, yr), 7)
    J = (out[0] ** 2).sum() + out[1].sum()
    ref = torch.autograd.grad(J, [xr, yr, a, b])
    xs, ys = x0.clone().requires_grad_(), y0.clone().requires_grad_()
    calls.clear()
    out2 = checkpointed_rollout(step, (xs, ys), 7, every=3, params=(a, fresh_w(), 1.5))
    assert not any(calls)  # the forward keeps no graph inside the segments
    assert torch.equal(out[0].detach(), out2[0].detach()) and torch.equal(out[1].detach(), out2[1].detach())
    J2 = (out2[0] ** 2).sum() + out2[1].sum()
    got = torch.autograd.grad(J2, [xs, ys, a, b])
    assert calls.count(True) == 7  # the backward recomputes every step once, with a graph
    for g, r in zip(got, ref):
        torch.testing.assert_close(g, r, rtol=1e-13, atol=1e-13)


def test_segments_without_params_lose_nothing_they_are_given():
    """Only the state and the listed params get gradients: a parameter the
    step reads but that is not listed gets none (the caller's contract)."""
    x0, y0, a, b, fresh_w, step, _ = _problem()
    xs = x0.clone().requires_grad_()
    out = checkpointed_rollout(step, (xs, y0.clone()), 4, every=2, params=(fresh_w(),))
    ga, gb, gx = torch.autograd.grad(out[0].sum(), [a, b, xs], allow_unused=True)
    assert ga is None and gb is not None and gx is not None
