"""Dense prefix-gain evaluation for the `effect` mode (realworld/exact.py:8-17, 65-82).

f(S_r) for every prefix r from the eigenvalues of the dense matrix M_S = M0 + sum_{e in S} b b^T,
computed with torch.linalg.eigvalsh in FP64 on the GPU (rocSOLVER); like the rest of the product
path it has no CPU fallback.  Used when n <= dense_max, as run_real.evaluate does
(run_real.py:52-56, 91-97).
"""
import numpy as np
import torch


def _add(M, a, b):
    """realworld/exact.py:8-9."""
    M[a, a] += 1
    M[b, b] += 1
    M[a, b] -= 1
    M[b, a] -= 1


def frac_value(lam, s, obj):
    """realworld/exact.py:12-17: sum lam^-s, sum 1/lam, or -sum log lam (log-det decrease)."""
    if obj == 'frac':
        return float(torch.sum(lam ** (-s)))
    if obj == 'er':
        return float(torch.sum(1.0 / lam))
    return -float(torch.sum(torch.log(lam)))


def _dense(M0):
    if not torch.cuda.is_available():
        raise RuntimeError('the dense evaluator needs a ROCm GPU (torch.cuda.is_available() is False); '
                           'there is no CPU fallback')
    return torch.as_tensor(np.asarray(M0, dtype=np.float64), device='cuda').clone()


def eval_prefix(M0, cand, seq, s, obj='frac'):
    """Exact gains f(S_r) for every prefix r (realworld/exact.py:65-71)."""
    M = _dense(M0)
    base = frac_value(torch.linalg.eigvalsh(M), s, obj)
    out = []
    for e in seq:
        _add(M, int(cand[e, 0]), int(cand[e, 1]))
        out.append(base - frac_value(torch.linalg.eigvalsh(M), s, obj))
    return np.array(out)


def eval_prefix_multi(M0, cand, seq, targets):
    """Prefix gains for several (obj, s) targets, one eigen-solve per prefix (exact.py:74-82)."""
    M = _dense(M0)
    lam = torch.linalg.eigvalsh(M)
    base = [frac_value(lam, s, o) for o, s in targets]
    out = [[] for _ in targets]
    for e in seq:
        _add(M, int(cand[e, 0]), int(cand[e, 1]))
        lam = torch.linalg.eigvalsh(M)
        for i, (o, s) in enumerate(targets):
            out[i].append(base[i] - frac_value(lam, s, o))
    return [np.array(v) for v in out]
