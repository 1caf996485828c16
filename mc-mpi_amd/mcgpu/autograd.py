"""torch.autograd binding of the transport: per-cell absorbed weight as a
differentiable function of the per-cell cross sections (sig_a, sig_s).

Forward: mcgpu_forward_cells on a copy of the source bank.  Backward:
mcgpu_adjoint_cells replays every history from the saved initial bank (seed
replay, no stored paths) with g = grad_output as target weights, giving the
pathwise derivative for sig_a and the score-function derivative for sig_s.
If the target weights are known when the forward runs (target_hint), the
forward also stores each history's weighted tally and the backward replays
each history once instead of twice.
"""
import torch

from . import check_bank, check_tensor, make_options


class SlabProblem:
    """Geometry and source of one slab: the bank (dict of device tensors) is
    the initial state of every history and is never modified."""

    def __init__(self, ctx, bank, min_index, max_index, dx, particle_min_weight, options=None):
        self.ctx = ctx
        self.bank = bank
        self.min_index, self.max_index = min_index, max_index
        self.dx, self.min_w = dx, particle_min_weight
        self.options = options if options is not None else make_options()
        if max_index <= min_index:
            raise ValueError("max_index must exceed min_index")
        check_bank(bank, "bank", device=getattr(ctx, "device", None))

    @property
    def m(self):
        return self.max_index - self.min_index

    @property
    def n(self):
        return self.bank["x"].numel()


class SlabTally(torch.autograd.Function):
    """tally = SlabTally.apply(sig_a, sig_s, problem, target_hint)"""

    @staticmethod
    def forward(fctx, sig_a, sig_s, problem, target_hint=None):
        # The library reads m doubles of each and the backward writes m
        # doubles of gradient: anything else is rejected before any launch.
        d = getattr(problem.ctx, "device", None)
        check_tensor(sig_a, "sig_a", torch.float64, problem.m, d)
        check_tensor(sig_s, "sig_s", torch.float64, problem.m, d)
        if target_hint is not None:
            check_tensor(target_hint, "target_hint", torch.float64, problem.m, d)
        check_bank(problem.bank, "problem.bank", device=d)
        sig_a = sig_a.detach()
        sig_s = sig_s.detach()
        work = {k: v.clone() for k, v in problem.bank.items()}
        tally = torch.zeros(problem.m, dtype=torch.float64, device=sig_a.device)
        score = None
        g = None
        if target_hint is not None:
            g = target_hint.detach()
            score = torch.empty(problem.n, dtype=torch.float64, device=sig_a.device)
        problem.ctx.forward_cells(work, sig_a, sig_s, tally, problem.min_index,
                                  problem.max_index, problem.dx, problem.min_w, g=g,
                                  history_score=score, options=problem.options)
        fctx.problem = problem
        fctx.hint = g
        fctx.score = score
        fctx.save_for_backward(sig_a, sig_s)
        return tally

    @staticmethod
    def backward(fctx, grad_tally):
        sig_a, sig_s = fctx.saved_tensors
        p = fctx.problem
        g = grad_tally.detach().to(torch.float64).contiguous()
        check_tensor(g, "grad_output", torch.float64, p.m, getattr(p.ctx, "device", None))
        score = None
        if fctx.hint is not None and torch.equal(g, fctx.hint):
            score = fctx.score
        ga = torch.zeros_like(sig_a)
        gs = torch.zeros_like(sig_s)
        p.ctx.adjoint_cells(p.bank, sig_a, sig_s, p.min_index, p.max_index, p.dx, p.min_w, g,
                            ga, gs, history_score=score, options=p.options)
        return ga, gs, None, None


def slab_tally(sig_a, sig_s, problem, target_hint=None):
    return SlabTally.apply(sig_a, sig_s, problem, target_hint)
