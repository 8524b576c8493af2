"""The synthetic workloads the core is measured and checked on (host-side inputs).

    c1  BA n=2,000, m0=5, seed 0; sigma=0.1, s=0.5; q=32 Rademacher; 16 poles; N=5,000 uniform
        non-edges; k=10, B=1                                              (CPU-runnable, FP64)
    c2  Chung-Lu n=100,000, gamma=2.5, mean degree 20, seed 1; sigma=0.05, s in {.25,.5,.75};
        q=64; 24 poles; N=1,000,000; k=20, B=8                            (FP64, then FP32)
    c3  Chung-Lu n=1,000,000, gamma=2.3, mean degree 16, weights U[0.5,1.5], seed 2; sigma=0.05,
        s=0.5, q=64, 24 poles; 2,000,000 candidate non-edges; k=20, B=16
    c4  64 R-MAT graphs (0.57,0.19,0.19,0.05), n log-uniform in [1e4, 5e5], edge factor 8,
        seeds 100-163; sigma=0.05, q=32, 16 poles, k=10, B=4, N=2,000
    c5  Chung-Lu n=4,000,000, gamma=2.2, m=34.7M, seed 3; sigma=0.02, s=0.5; q=128; 32 poles;
        N=10,000,000; k=50, B=16

Poles: quadrature.fixed_rule(s, p, sigma, Lam(k)).  Sketches: counter-hash Rademacher with the
config seed (polar_gpu.sketches kind='rademacher').  `scale` < 1 shrinks n, m and N for tests.
"""
import math

import numpy as np

from . import graphs
from .quadrature import fixed_rule

TOL = 1e-7

SPECS = {
    'c1': dict(kind='ba', n=2000, m0=5, seed=0, sigma=0.1, s=(0.5,), q=32, p=16, N=5000, k=10, B=1),
    'c2': dict(kind='chung_lu', n=100_000, gamma=2.5, mean_deg=20.0, seed=1, sigma=0.05,
               s=(0.25, 0.5, 0.75), q=64, p=24, N=1_000_000, k=20, B=8),
    'c3': dict(kind='chung_lu', n=1_000_000, gamma=2.3, mean_deg=16.0, seed=2, sigma=0.05, s=(0.5,),
               q=64, p=24, N=2_000_000, k=20, B=16, weighted=True),
    'c5': dict(kind='chung_lu', n=4_000_000, gamma=2.2, mean_deg=17.35, m=34_700_000, seed=3,
               sigma=0.02, s=(0.5,), q=128, p=32, N=10_000_000, k=50, B=16),
}


def rmat_spec(i):
    """Graph i (0..63) of c4: n log-uniform in [1e4, 5e5] drawn from seed 100+i."""
    rng = np.random.default_rng(100 + i)
    n = int(round(math.exp(rng.uniform(math.log(1e4), math.log(5e5)))))
    s = float(rng.choice([0.25, 0.5, 0.75]))
    return dict(kind='rmat', n=n, seed=100 + i, sigma=0.05, s=(s,), q=32, p=16, N=2000, k=10, B=4)


def make(name, scale=1.0, s=None):
    """Host inputs of a workload: dict(n, edges, weights, sigma, cand, s, q, p, k, B, tau, om, seed).
    Sketches are generated separately (polar_gpu.sketches(..., kind='rademacher', seed=seed))."""
    if name.startswith('c4_'):
        spec = rmat_spec(int(name[3:]))
    else:
        spec = dict(SPECS[name])
    n = max(64, int(round(spec['n'] * scale)))
    seed = spec['seed']
    weights = None
    if spec['kind'] == 'ba':
        edges = graphs.barabasi_albert(n, spec['m0'], seed)
    elif spec['kind'] == 'chung_lu':
        m = spec.get('m')
        m = None if m is None else int(round(m * scale))
        edges = graphs.chung_lu(n, spec['gamma'], spec['mean_deg'], seed, m=m)
        if spec.get('weighted'):
            weights = np.random.default_rng(seed + 1_000_003).uniform(0.5, 1.5, len(edges))
    elif spec['kind'] == 'rmat':
        edges = graphs.rmat(n, 0.57, 0.19, 0.19, 0.05, 8, seed)
    else:
        raise ValueError(spec['kind'])
    N = max(16, int(round(spec['N'] * scale)))
    max_pairs = n * (n - 1) // 2 - len(edges)
    N = min(N, max_pairs // 2)
    if N <= 20000:
        cand = graphs.candidates(n, edges, N, seed)
    else:
        cand = graphs.candidates_fast(n, edges, N, seed)
    deg = np.bincount(edges[:, 0], weights, n) + np.bincount(edges[:, 1], weights, n)
    sigma = spec['sigma']
    k = spec['k']
    Lam = sigma + 2 * float(deg.max()) + 4 * k
    ss = spec['s'][0] if s is None else s
    tau, om, _ = fixed_rule(ss, spec['p'], sigma, Lam)
    return dict(name=name, n=n, edges=edges, weights=weights, sigma=sigma, cand=cand, s=ss,
                s_all=spec['s'], q=spec['q'], p=spec['p'], k=k, B=spec['B'], tau=tau, om=om,
                seed=seed, tol=TOL, N=N, m=len(edges), dmax=float(deg.max()))
