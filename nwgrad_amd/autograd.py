"""Soft alignment scores as a differentiable PyTorch function.

``soft_scores(theta, batch)`` returns the per-pair soft scores T*log Z of a ``grad_mode="soft"``
batch as a float64 CUDA tensor whose gradient with respect to ``theta`` is the adjoint of the
soft pass: for an upstream gradient w (one weight per pair), d(sum_p w_p s_p)/d(theta) is the
weighted sum of the per-pair gradient rows -- the reference's ``weighted_grad(w)``
(batch_engine.hpp:702-718), computed on the device by the library's blocked reduction.

``theta`` is the parameter vector of ``K * (n_alpha**2 + 4)`` doubles for K parameter sets: the K
substitution matrices row-major, then K rows of gap_open_a, gap_extend_a, gap_open_b,
gap_extend_b (the layout of the reference's gradient rows, ``grads()``; K = 1 for untracked
params).  Everything stays on the device: the parameters are
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
    """``AlignParams`` (K parameter sets) as the float64 vector ``theta``: the K matrices, then K
    rows of the four gaps."""
    mats = np.asarray(params.matrices, dtype=np.float64).ravel()
    gaps = np.asarray(params.gaps, dtype=np.float64).ravel()
    return torch.tensor(np.concatenate([mats, gaps]), dtype=torch.float64, device=device)


def align_params(theta: torch.Tensor, alphabet: str, num_sets: int = 1) -> AlignParams:
    """The inverse of :func:`params_vector`."""
    t = theta.detach().to("cpu", torch.float64).numpy()
    na, K = len(alphabet), int(num_sets)
    if t.size != K * (na * na + 4):
        raise ValueError(f"nwgrad_amd.autograd: theta has {t.size} values, {K} sets need {K * (na * na + 4)}")
    return AlignParams.from_arrays(t[:K * na * na].reshape(K, na, na).copy(), t[K * na * na:].reshape(K, 4).copy(),
                                   alphabet=alphabet)


def _bind_stream(batch):
    batch._set_stream(torch.cuda.current_stream().cuda_stream)


class SoftScores(torch.autograd.Function):
    """scores = soft pass(theta); d scores / d theta through the weighted-gradient reduction."""

    @staticmethod
    def forward(ctx, theta, batch):
        if batch.grad_mode != "soft":
            raise ValueError("nwgrad_amd.autograd: the batch must have grad_mode=\"soft\"")
        na, K = len(batch.alphabet), batch.num_sets
        if theta.dtype != torch.float64 or theta.numel() != K * (na * na + 4) or not theta.is_cuda:
            raise ValueError(f"nwgrad_amd.autograd: theta must be a float64 CUDA tensor of {K * (na * na + 4)} values "
                             f"({K} parameter set(s))")
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

    This call only enqueues work, so the float32 pass's refusals (pairs whose lost mass could
    exceed 2^-20) are not rerun inside it.  A refused pair's gradient row is left out of the
    backward's weighted sum on the device -- a step never applies a float32 gradient that the
    pass refused -- and the batch records the refusal.  Call ``finish(batch)`` after the step
    (or after a run of captured steps): it reruns the last pass's refusals in double and raises
    if any step since the previous ``finish`` had a refusal.
    """
    return SoftScores.apply(theta, batch)


class SoftRefusalError(RuntimeError):
    """Steps ran with float32 refusals masked out of their gradients."""


def finish(batch) -> int:
    """Synchronise, rerun in double the pairs the last float32 pass refused (the number rerun is
    returned), and check the batch's refusal record: if any pass since the previous ``finish``
    refused a pair, those steps applied gradients without it, and SoftRefusalError says from
    which step on."""
    reruns = int(batch._soft_finish())
    refused, first, passes = batch._soft_refusals(True)
    if refused:
        raise SoftRefusalError(
            f"nwgrad_amd.autograd: the float32 soft pass refused {refused} pair-pass(es), first in pass {first} "
            f"of {passes}; those pairs' gradients were left out of their steps (the last pass's refusals were "
            f"rerun in double: {reruns}).  Use soft_precision=\"double\" or a higher temperature.")
    return reruns


def margin_loss(scores: torch.Tensor, n_pos: int, margin: float) -> torch.Tensor:
    """Contrastive margin loss over paired scores, pair k (a homolog) against pair n_pos + k
    (a decoy): mean(relu(margin - (s_pos - s_neg)))."""
    return torch.relu(margin - (scores[:n_pos] - scores[n_pos:2 * n_pos])).mean()
