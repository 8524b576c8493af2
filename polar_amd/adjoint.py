"""Adjoint form: the sketched fractional trace and its gradient with respect to edge weights.

    T(w) = sum_j om_j tr(Z^T (L(w) + sigma I + tau_j I)^{-1} Z)        (~ Tr((L+sigma I)^{-s}))

For a pair e = (u, v) with weight w_e, dM/dw_e = b_e b_e^T (b_e = e_u - e_v), so

    dT/dw_e = - sum_j om_j || X_j[u, :] - X_j[v, :] ||^2 ,   X_j = (M + tau_j)^{-1} Z.

The system is symmetric, so the adjoint solve is the forward solve itself: the gradient of every
existing edge and of every candidate non-edge (weight 0) comes from the forward's shifted
solution blocks, gathered by the `polar_pair_wsum` kernel.  The gradient is a sum over sketch
columns, so the columns are processed in chunks and each chunk's shifted blocks are consumed
right after its solve (bounded memory whatever n*q*poles is).  -dT/dw_e at w_e = 0 is the
first-order proxy of the reference's h cache (realworld/polar_gpu.py:205, summed over poles with
the quadrature weights); `gradient_shortlist` ranks candidates by it.

The trace value uses the multi-shift CG's trace tracker (w^T x = sum_t a_t prod b, FP64).
"""
import math

import numpy as np
import torch

from . import _lib as L
from . import polar_gpu as pg


def sketched_trace(n, pairs, w, sigma, tau, om, Z, tol=1e-10, precision='fp64', want_grad=True,
                   width=None):
    """(T, grad) for pairs (E, 2) with weights w (E,) (zeros allowed: candidate non-edges).
    Z: (n, q) sketch (numpy or device tensor).  Returns T (float) and grad (device float64 (E,),
    dT/dw) or None."""
    lib = L.require_gpu()
    dev = torch.device('cuda', torch.cuda.current_device())
    pairs = np.asarray(pairs, dtype=np.int64)
    w_h = w.detach().double().cpu().numpy() if torch.is_tensor(w) else np.asarray(w, dtype=np.float64)
    keep = w_h != 0
    G = pg.Graph(n, pairs[keep], sigma, weights=w_h[keep])
    op = pg.Op(G, precision=precision)
    dtype = op.dtype
    Zd = torch.as_tensor(Z, device=dev).to(dtype)
    q = Zd.shape[1]
    tau = np.asarray(tau, dtype=np.float64)
    om = np.asarray(om, dtype=np.float64)
    p = len(tau)
    om_d = torch.from_numpy(om).to(dev)
    active = torch.arange(n, dtype=torch.int32, device=dev)
    ea = torch.from_numpy(pairs[:, 0].astype(np.int32)).to(dev)
    eb = torch.from_numpy(pairs[:, 1].astype(np.int32)).to(dev)
    E = len(pairs)
    grad = torch.zeros(E, dtype=torch.float64, device=dev) if want_grad else None
    if width is None:
        width = pg.bank_width(n, p, n, q, dtype)
    terms = []
    for lo in range(0, q, width):
        hi = min(q, lo + width)
        X, _, _, te = pg.msolve(op, Zd[:, lo:hi], tau, active, tol, trace=True)
        terms.append((om[:, None] * te.cpu().numpy()).ravel())
        if want_grad:
            L.check(lib.polar_pair_wsum(L.dtype_code(dtype), L.ptr(X), X.stride(0), X.stride(1), p, hi - lo,
                                        L.ptr(ea), L.ptr(eb), E, L.ptr(om_d), L.ptr(grad), 1 if lo else 0,
                                        L.stream_ptr()), 'pair_wsum')
        del X
    T = math.fsum(np.concatenate(terms))
    return T, (-grad if want_grad else None)


class SketchedTrace(torch.autograd.Function):
    """torch.autograd.Function: T = SketchedTrace.apply(w, n, pairs, sigma, tau, om, Z, tol,
    precision); T.backward() gives dT/dw for every pair (edges and zero-weight candidates)."""

    @staticmethod
    def forward(ctx, w, n, pairs, sigma, tau, om, Z, tol=1e-10, precision='fp64'):
        T, g = sketched_trace(n, pairs, w, sigma, tau, om, Z, tol=tol, precision=precision,
                              want_grad=ctx.needs_input_grad[0])
        if g is not None:
            ctx.save_for_backward(g.to(device=w.device, dtype=w.dtype))
        return torch.tensor(T, dtype=torch.float64, device=w.device)

    @staticmethod
    def backward(ctx, gout):
        (g,) = ctx.saved_tensors
        return (gout.to(device=g.device, dtype=g.dtype) * g,) + (None,) * 8


def gradient_shortlist(grad_cand, B):
    """Indices of the B candidates with the largest first-order gain -dT/dw (ties: lower index)."""
    g = -grad_cand.detach().cpu().numpy()
    return np.argsort(-g, kind='stable')[:B].tolist()
