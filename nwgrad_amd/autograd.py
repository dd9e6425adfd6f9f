"""Soft alignment scores as a differentiable PyTorch function.

``soft_scores(theta, batch)`` returns the per-pair soft scores T*log Z of a ``grad_mode="soft"``
batch as a float64 CUDA tensor whose gradient with respect to ``theta`` is the adjoint of the
soft pass: for an upstream gradient w (one weight per pair), d(sum_p w_p s_p)/d(theta) is the
weighted sum of the per-pair gradient rows -- the reference's ``weighted_grad(w)``
(batch_engine.hpp:702-718), computed on the device by the library's blocked reduction.

``theta`` is the parameter vector of ``n_alpha**2 + 4`` doubles: the substitution matrix
row-major, then gap_open_a, gap_extend_a, gap_open_b, gap_extend_b (the layout of the
reference's gradient rows, ``grads()``).  Everything stays on the device: the parameters are
copied into the batch, the soft pass and the reduction run on torch's current stream, so a
training step built from these calls can be captured in a CUDA (HIP) graph once the batch has
run one uncaptured step (which sizes its device scratch).

PyTorch is the wrapper here only; the forward-backward itself is the library's HIP kernel.
"""
from __future__ import annotations

import numpy as np
import torch

from . import AlignParams


def params_vector(params: AlignParams, device="cuda") -> torch.Tensor:
    """``AlignParams`` (one parameter set) as the float64 vector ``theta``."""
    d = params.to_dict()
    mat = np.asarray(d["matrix"], dtype=np.float64).ravel()
    gaps = np.array([d["gap_open_a"], d["gap_extend_a"], d["gap_open_b"], d["gap_extend_b"]], dtype=np.float64)
    return torch.tensor(np.concatenate([mat, gaps]), dtype=torch.float64, device=device)


def align_params(theta: torch.Tensor, alphabet: str) -> AlignParams:
    """The inverse of :func:`params_vector`."""
    t = theta.detach().to("cpu", torch.float64).numpy()
    na = len(alphabet)
    return AlignParams(t[:na * na].reshape(na, na).copy(), alphabet=alphabet, gap_open_a=float(t[na * na]),
                       gap_extend_a=float(t[na * na + 1]), gap_open_b=float(t[na * na + 2]),
                       gap_extend_b=float(t[na * na + 3]))


def _bind_stream(batch):
    batch._set_stream(torch.cuda.current_stream().cuda_stream)


class SoftScores(torch.autograd.Function):
    """scores = soft pass(theta); d scores / d theta through the weighted-gradient reduction."""

    @staticmethod
    def forward(ctx, theta, batch):
        if batch.grad_mode != "soft":
            raise ValueError("nwgrad_amd.autograd: the batch must have grad_mode=\"soft\"")
        na = len(batch.alphabet)
        if theta.dtype != torch.float64 or theta.numel() != na * na + 4 or not theta.is_cuda:
            raise ValueError(f"nwgrad_amd.autograd: theta must be a float64 CUDA tensor of {na * na + 4} values")
        theta = theta.contiguous()
        _bind_stream(batch)
        batch._set_params_dev(theta.data_ptr())
        batch._run_async()
        scores = torch.empty(len(batch), dtype=torch.float64, device=theta.device)
        batch._scores_copy_dev(scores.data_ptr())
        ctx.batch = batch
        ctx.n_theta = theta.numel()
        return scores

    @staticmethod
    def backward(ctx, grad_scores):
        batch = ctx.batch
        out = torch.empty(ctx.n_theta, dtype=torch.float64, device=grad_scores.device)
        w = grad_scores.to(torch.float64).contiguous()
        _bind_stream(batch)
        batch._weighted_grad_dev(w.data_ptr(), out.data_ptr())
        return out, None


def soft_scores(theta: torch.Tensor, batch) -> torch.Tensor:
    """Differentiable per-pair soft scores of ``batch`` at parameters ``theta``.

    The float32 pass's refusals (pairs whose lost mass could exceed 2^-20) are not rerun
    inside this call, which enqueues work only; call ``finish(batch)`` after the step (or
    after a run of captured steps) to rerun them in double and raise if any pair fits neither.
    """
    return SoftScores.apply(theta, batch)


def finish(batch) -> int:
    """Synchronise, rerun in double the pairs the last float32 pass refused; the number rerun."""
    return int(batch._soft_finish())


def margin_loss(scores: torch.Tensor, n_pos: int, margin: float) -> torch.Tensor:
    """Contrastive margin loss over paired scores, pair k (a homolog) against pair n_pos + k
    (a decoy): mean(relu(margin - (s_pos - s_neg)))."""
    return torch.relu(margin - (scores[:n_pos] - scores[n_pos:2 * n_pos])).mean()
