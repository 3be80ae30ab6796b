"""torch.autograd wrappers of the elastic and heat misfits and their
hand-written adjoints.

Used for checking only (finite differences, autograd plumbing): the forward
and the gradient both run through the C ABI (b2r_el_gradient,
b2r_heat_gradient); torch only carries the arrays.  The elastic medium is
parametrised by (vp, vs, rho) with
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


class HeatMisfit(torch.autograd.Function):
    """J(u0, k) = 1/2 sum_owned (u_N - target)^2 for nsteps F32 heat steps of
    delta-D on a Context whose radius is D; target is a field of that context."""

    @staticmethod
    def forward(ctx, u0, k, context, target, D, nsteps, ckpt):
        from . import _abi as A
        kk = [float(v) for v in k.detach().cpu().numpy()]
        p = A.heat_params(D, A.HEAT_F32, kk)
        fu, fg = context.field(), context.field()
        try:
            context.upload(fu, u0.detach().cpu().numpy().astype(np.float32))
            J, gk = context.heat_gradient(fu, target, p, nsteps, ckpt, fg)
            g = context.download(fg)
        finally:
            context.free(fu)
            context.free(fg)
        ctx.save_for_backward(torch.from_numpy(g.astype(np.float64)), torch.from_numpy(gk))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        g, gk = ctx.saved_tensors
        return gJ * g, gJ * gk, None, None, None, None, None


def heat_misfit(u0, k, context, target, D, nsteps, ckpt=50):
    return HeatMisfit.apply(u0, k, context, target, D, nsteps, ckpt)
