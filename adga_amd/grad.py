"""Gradients of chi_m(q, omega = 0) and Sigma_DGA(k, nu) with respect to the model parameters (t, t', U)
behind a torch.autograd.Function.

The forward pass runs the hot path on the GPU for the workload with parameters theta = (t, t', U) (every
(omega, q) point, complex128) and returns

* chi_m(q, omega = 0) for every q of the list: the reference's final magnetic susceptibility, non-local
  part + non-local bubble + local chi + local bubble (src/main.F90:935-936, 1124-1140), (n^2, n^2, nqp);
* Sigma_DGA(k, nu) = Sigma_nl + Sigma_DMFT on the diagonal (add_siw_dmft, src/eom_module.f90:237-251),
  (n, n, 2*iwfmax_small, nkp_eom).

The backward pass keeps nothing from the forward pass but theta: it runs the path again in its
tangent-linear mode (adga_run_points_tangent: the bubble formed again with one factor differentiated at a
time, each direction's derivative solved with the forward LU factors) along the three parameter directions
of synth.tangents, and contracts the derivatives with the incoming cotangents, grad_theta_p =
Re sum conj(g) . d(output)/d(theta_p), the exact vector-Jacobian product.  The U dependence is the synthetic
model's: the Hubbard-I-like self-energy, mu = U/2 and the Kanamori U, U~ (J fixed).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import api, synth

__all__ = ["ChiSigma", "forward_arrays", "tangent_arrays", "with_params"]


def with_params(wl, theta):
    t, tp, U = (float(x) for x in theta)
    return dataclasses.replace(wl, t=t, tp=tp, U=U)


def _points(wl):
    return [(w, q) for w in range(-wl.iwbmax_small, wl.iwbmax_small + 1) for q in range(wl.nqp)]


def _assemble(out, inp, points, tangent=False):
    wl = inp.wl
    n, iwf = wl.n, wl.iwfmax_small
    d = "d" if tangent else ""
    sel = [p for p, (w, _) in enumerate(points) if w == 0]
    chi = out[d + "chi_m"][:, :, sel] + out[d + "bubble_nl"][:, :, sel]
    loc = out[d + "chi_loc_m"][0] + out[d + "bubble_loc"][0]
    chi = chi + (loc[:, :, None, :] if tangent else loc[:, :, None])
    sig = out[d + "sigma_nl"].copy()
    off = inp.iwmax
    if tangent:
        dsiw = synth.tangents(inp).dsiw
        for a in range(n):
            sig[a, a] += dsiw[off - iwf:off + iwf, a, :][:, None, :]
    else:
        for a in range(n):
            sig[a, a] += inp.siw[off - iwf:off + iwf, a][:, None]
    return chi, sig


def forward_arrays(wl, points=None, batch=None):
    """(chi_m(q, omega = 0), Sigma_DGA) of workload wl on the GPU, complex128."""
    inp = synth.build_inputs(wl)
    points = _points(wl) if points is None else points
    with api.Context(inp, batch=batch) as ctx:
        out = api.run_loop(ctx, points, lambda w, c: synth.chi_loc(inp, w, c), "fp64")
    if (out["status"] != 0).any():
        raise api.AdgaError(int(out["status"].max()), "a point failed in the forward pass")
    return _assemble(out, inp, points)


def tangent_arrays(wl, points=None, batch=None, eom=None, bubble=None):
    """d(chi_m(q, omega = 0))/d theta and d Sigma_DGA / d theta, the last axis over (t, t', U); `eom` and
    `bubble` the context's modes (0 direct sums, 1 k-grid convolutions; defaults 1 and 2, see Context)."""
    inp = synth.build_inputs(wl)
    tan = synth.tangents(inp)
    points = _points(wl) if points is None else points
    with api.Context(inp, batch=batch, eom=eom, bubble=bubble) as ctx:
        ctx.set_tangents(tan)
        out = api.run_loop_tangent(ctx, points, lambda w, c: synth.chi_loc(inp, w, c),
                                   lambda w, c: synth.dchi_loc(inp, tan, w))
    if (out["status"] != 0).any():
        raise api.AdgaError(int(out["status"].max()), "a point failed in the tangent pass")
    return _assemble(out, inp, points, tangent=True)


class ChiSigma(torch.autograd.Function):
    """theta (3,) float64 tensor (t, t', U) -> (chi_m(q, omega = 0), Sigma_DGA) as complex128 tensors on
    theta's device; `workload` gives everything else (a synth.Workload), `points` optionally a subset of
    the (omega, q) points (the self-energy then sums over that subset)."""

    @staticmethod
    def forward(ctx, theta, workload, points=None):
        wl = with_params(workload, theta.detach().cpu().numpy())
        chi, sig = forward_arrays(wl, points)
        ctx.wl = wl
        ctx.points = points
        ctx.save_for_backward(theta)
        return (torch.from_numpy(np.ascontiguousarray(chi)).to(theta.device),
                torch.from_numpy(np.ascontiguousarray(sig)).to(theta.device))

    @staticmethod
    def backward(ctx, g_chi, g_sig):
        (theta,) = ctx.saved_tensors
        dchi, dsig = tangent_arrays(ctx.wl, ctx.points)
        gc = g_chi.detach().cpu().numpy()
        gs = g_sig.detach().cpu().numpy()
        grad = [float(np.real(np.vdot(gc, dchi[..., p])) + np.real(np.vdot(gs, dsig[..., p]))) for p in range(3)]
        return torch.tensor(grad, dtype=theta.dtype, device=theta.device), None, None
