"""Stieltjes quadrature rules for x^-s (host, FP64).

* composite(s, a, z, eps): the positive composite rule of realworld/polar_gpu.py:18-36
  (Gauss-Jacobi u^-s panel on [0,a], K Gauss-Legendre dyadic panels, Gauss-Jacobi u^(s-1) tail
  mapped by C/v); p=(K+2)r poles with r=ceil(log(256/eps)/log 4), K=ceil(log2(z/a)).
* fixed_rule(s, p, lo, hi): a fixed-count p-pole rule (the Gauss-Jacobi rule of
  synthetic/scalability/benchmark.py:20-25 with t=(1-y)/(1+y)), rescaled by c=sqrt(lo*hi) so it
  covers the spectrum [lo, hi]; used where a pole count (16/24/32) is prescribed.
* poles(obj, ...): polar_gpu.py:39-43.
"""
import math

import numpy as np


def _golub_welsch(diag, off):
    x, Q = np.linalg.eigh(np.diag(diag) + np.diag(off, 1) + np.diag(off, -1))
    return x, Q[0] ** 2


def _jacobi01(beta, r):
    """Gauss rule on [0,1] for weight u^beta from the (0,beta) Jacobi recurrence."""
    i = np.arange(1, r, dtype=float)
    diag = np.empty(r)
    diag[0] = beta / (beta + 2)
    diag[1:] = beta * beta / ((2 * i + beta) * (2 * i + beta + 2))
    off = 2 / (2 * i + beta) * np.sqrt(i * i * (i + beta) ** 2 / ((2 * i + beta - 1) * (2 * i + beta + 1)))
    x, w0 = _golub_welsch(diag, off)
    return (1 + x) / 2, w0 / (beta + 1)


def composite(s, a, z, eps):
    r = int(np.ceil(np.log(256 / eps) / np.log(4)))
    K = int(np.ceil(np.log2(z / a)))
    cs = np.sin(np.pi * s) / np.pi
    lo, wlo = _jacobi01(-s, r)
    hi, whi = _jacobi01(s - 1, r)
    mid, wmid = np.polynomial.legendre.leggauss(r)
    nodes = [a * lo]
    weights = [cs * a ** (1 - s) * wlo]
    for ell in range(K):
        left = a * 2.0 ** ell
        t = left * (3 + mid) / 2
        nodes.append(t)
        weights.append(cs * left / 2 * wmid * t ** (-s))
    C = a * 2.0 ** K
    nodes.append(C / hi)
    weights.append(cs * C ** (1 - s) * whi / hi)
    tau = np.concatenate(nodes)
    om = np.concatenate(weights)
    order = np.argsort(tau)
    return tau[order], om[order], 256 * 4.0 ** (-r)


def fixed_rule(s, p, lo, hi):
    i = np.arange(p)
    diag = (1 - 2 * s) / (4 * i * i - 1.0)
    j = np.arange(1, p)
    off = np.sqrt((j - s) * (j + s - 1)) / (2 * j - 1)
    off[0] *= np.sqrt(2)
    y, w0 = _golub_welsch(diag, off)
    t = ((1 - y) / (1 + y))[::-1]
    w = (2 * w0 / (1 + y))[::-1]
    c = math.sqrt(lo * hi)
    return c * t, c ** (1 - s) * w, float('nan')


def poles(obj, s, sigma, Lam, eps):
    if obj in ('er', 'logdet'):
        return np.zeros(1), np.ones(1), 0.0
    return composite(s, sigma, Lam, eps)
