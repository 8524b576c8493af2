"""SLQ-Greedy baseline on the MI355X core (realworld/slq.py:10-58).

Each round estimates every live candidate's trace difference Tr(f(M_S)) - Tr(f(M_S + b b^T)),
f(x) = x^-s, by Hutchinson probes shared by the base matrix and every candidate (common random
numbers), each z^T f(A) z by m-step Lanczos quadrature; the argmax is selected.  Candidates and
probes are batched as columns: column c = (candidate, probe) applies the current operator (our
SpMM, libpolar_amd.so) plus its own candidate's rank-1 term.  The three-term recurrences run in
FP64 on the device; the m x m tridiagonal eigenproblems of all columns are one batched eigh.
"""
import time

import numpy as np
import torch

from . import _lib as L
from . import polar_gpu as pg


class _Apply:
    """Y = M_S X for the current added edges, on the device, into a reused output buffer."""

    def __init__(self, op):
        self.op = op
        self.lib = L.require_gpu()
        self.ws = None
        self.Y = None

    def __call__(self, X):
        n, C = X.shape
        if self.Y is None or self.Y.shape != X.shape:
            self.Y = torch.empty_like(X)
        s = self.op.c_op(X.dtype)
        need = max(self.lib.polar_spmm_ws_bytes(s, min(C, 128)), 256)
        if self.ws is None or self.ws.numel() < need:
            self.ws = torch.empty(need, dtype=torch.uint8, device=X.device)
        for lo in range(0, C, 128):
            r = min(128, C - lo)
            L.check(self.lib.polar_spmm(s, 0.0, L.ptr(X[:, lo:]), X.stride(0), L.ptr(self.Y[:, lo:]), self.Y.stride(0),
                                        r, L.ptr(self.ws), self.ws.numel(), L.stream_ptr()), 'polar_spmm')
        return self.Y


def lanczos_quad(apply, Z, m, f):
    """z_c^T f(A_c) z_c for every column of Z (n x C, FP64 device) by m-step Lanczos quadrature
    (slq.py:10-21); apply(V) returns A_c v_c column by column."""
    n, C = Z.shape
    nz = torch.sqrt(torch.sum(Z * Z, 0))
    V = Z / nz
    Vp = torch.zeros_like(V)
    al = torch.zeros((m, C), dtype=Z.dtype, device=Z.device)
    be = torch.zeros((m, C), dtype=Z.dtype, device=Z.device)
    b = torch.zeros(C, dtype=Z.dtype, device=Z.device)
    for j in range(m):
        W = apply(V) - b * Vp
        a = torch.sum(W * V, 0)
        W = W - a * V
        b = torch.sqrt(torch.sum(W * W, 0))
        al[j] = a
        be[j] = b
        Vp = V
        V = W / torch.where(b > 0, b, torch.ones_like(b))
    T = torch.zeros((C, m, m), dtype=Z.dtype, device=Z.device)
    idx = torch.arange(m, device=Z.device)
    T[:, idx, idx] = al.T
    T[:, idx[:-1], idx[1:]] = be[:-1].T
    T[:, idx[1:], idx[:-1]] = be[:-1].T
    th, U = torch.linalg.eigh(T)
    return nz ** 2 * torch.sum(U[:, 0, :] ** 2 * f(torch.clamp(th, min=1e-300)), dim=1)


def slq_greedy(G, cand, k, s, probes=20, m=40, chunk=4000, seed=7, return_estimates=False):
    """slq.py:24-58 on the GPU.  Returns dict(sequence, seconds, spmv, probes, lanczos_steps) (and
    the per-round estimates with return_estimates=True)."""
    L.require_gpu()
    dev = pg._device()
    f = lambda x: x ** (-s)
    n = G.n
    cand = np.asarray(cand)
    rng = np.random.default_rng(seed)
    live = np.ones(len(cand), bool)
    seq, ests = [], []
    op = pg.Op(G, precision='fp64')
    base_apply = _Apply(op)
    t0 = time.perf_counter()
    spmv = 0
    for _ in range(k):
        Z = torch.from_numpy(rng.choice([-1.0, 1.0], size=(n, probes))).to(dev)
        base = float(torch.mean(lanczos_quad(base_apply, Z, m, f)))
        spmv += probes * m
        ids = np.flatnonzero(live)
        est = np.empty(len(ids))
        per = max(1, chunk // probes)
        cand_apply = _Apply(op)
        for lo in range(0, len(ids), per):
            sub = ids[lo:lo + per]
            c = len(sub)
            ea = torch.from_numpy(np.repeat(cand[sub, 0], probes)).to(dev)
            eb = torch.from_numpy(np.repeat(cand[sub, 1], probes)).to(dev)
            cols = torch.arange(c * probes, device=dev)

            def apply(X):
                Y = cand_apply(X)
                d = X[ea, cols] - X[eb, cols]
                Y[ea, cols] += d
                Y[eb, cols] -= d
                return Y
            q = lanczos_quad(apply, Z.repeat(1, c), m, f).reshape(c, probes)
            est[lo:lo + c] = (base - torch.mean(q, 1)).cpu().numpy()
            spmv += c * probes * m
        e = int(ids[np.argmax(est)])
        seq.append(e)
        ests.append(est)
        live[e] = False
        op.add(int(cand[e, 0]), int(cand[e, 1]))
    torch.cuda.synchronize()
    out = dict(sequence=seq, seconds=time.perf_counter() - t0, spmv=spmv, probes=probes, lanczos_steps=m)
    if return_estimates:
        out['estimates'] = ests
    return out
