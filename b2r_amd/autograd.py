"""torch.autograd wrapper of the elastic misfit and its hand-written adjoint.

Used for checking only (finite differences, autograd plumbing): the forward
and the gradient both run through the C ABI (b2r_el_gradient); torch only
carries the arrays.  The medium is parametrised by (vp, vs, rho) with
lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2, b = 1/rho, so

    dJ/dvp  = 2 rho vp g_lambda
    dJ/dvs  = -4 rho vs g_lambda + 2 rho vs g_mu
    dJ/drho = (vp^2 - 2 vs^2) g_lambda + vs^2 g_mu - g_b / rho^2
"""
from __future__ import annotations

import numpy as np
import torch

from . import workloads as W


class ElasticMisfit(torch.autograd.Function):
    """J(vp, vs, rho) = 1/2 sum_t,r (rec - obs)^2 for one ElasticShot whose
    sources and receivers are already set."""

    @staticmethod
    def forward(ctx, vp, vs, rho, shot, obs, nsteps, ckpt):
        vp_n, vs_n, rho_n = (t.detach().cpu().numpy().astype(np.float32) for t in (vp, vs, rho))
        lam, mu, b = W.lame(vp_n, vs_n, rho_n)
        shot.set_medium(lam, mu, b)
        J, (g_lam, g_mu, g_b) = shot.gradient(obs, nsteps, ckpt)
        vp64, vs64, rho64 = (a.astype(np.float64) for a in (vp_n, vs_n, rho_n))
        gl, gm, gb = (a.astype(np.float64) for a in (g_lam, g_mu, g_b))
        g_vp = 2 * rho64 * vp64 * gl
        g_vs = -4 * rho64 * vs64 * gl + 2 * rho64 * vs64 * gm
        g_rho = (vp64 ** 2 - 2 * vs64 ** 2) * gl + vs64 ** 2 * gm - gb / rho64 ** 2
        ctx.save_for_backward(*(torch.from_numpy(g) for g in (g_vp, g_vs, g_rho)))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        g_vp, g_vs, g_rho = ctx.saved_tensors
        return gJ * g_vp, gJ * g_vs, gJ * g_rho, None, None, None, None


def elastic_misfit(vp, vs, rho, shot, obs, nsteps, ckpt=50):
    return ElasticMisfit.apply(vp, vs, rho, shot, obs, nsteps, ckpt)
