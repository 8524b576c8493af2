"""CPU baseline of the greedy: the oracle (NumPy/SciPy FP64) timed on this host's cores.

TEST INFRASTRUCTURE ONLY -- the benchmark starts it as a child process (before the parent touches
the GPU) so that it can run on the host while the GPU legs run:

    python -m oracle.cpu_baseline --config c2 --s 0.5 --workers 12 --time-limit 60 [--wait]

With --wait it blocks until a line arrives on stdin, then runs oracle.polar() on the workload's
inputs (the same graph, candidate pool, Rademacher sketch and poles as the GPU path) under the
reference's time_limit convention (realworld/polar_gpu.py:215-216: the bank always completes,
round 0 always runs, later rounds start only while the elapsed time is below the limit) and
prints one JSON line: seconds, rounds, bank_seconds, the per-round seconds, column-iterations,
the sequence so far, and the cores it used.  Column blocks of every solve are split over
`workers` forked processes (columns are independent, so the results equal a one-process run).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--config', default='c2')
    ap.add_argument('--s', type=float, default=0.5)
    ap.add_argument('--workers', type=int, default=12)
    ap.add_argument('--time-limit', type=float, default=60.0)
    ap.add_argument('--scale', type=float, default=1.0)
    ap.add_argument('--wait', action='store_true')
    a = ap.parse_args()
    from oracle import polar_oracle as O
    from polar_amd import configs
    t = time.perf_counter()
    c = configs.make(a.config, a.scale, s=a.s)
    Go = O.Graph(c['n'], c['edges'], c['sigma'], weights=c['weights'])
    sk = O.sketches(Go, c['cand'], c['q'], c['seed'], kind='rademacher')
    prep = time.perf_counter() - t
    if a.wait:
        sys.stdin.readline()
    r = O.polar(Go, c['cand'], c['k'], c['s'], q=c['q'], B=c['B'], tol=c['tol'], sketch=sk,
                poles=(c['tau'], c['om']), width=2 * c['q'], time_limit=a.time_limit, workers=a.workers)
    rounds = r['rounds']
    per_round = (r['seconds'] - r['bank_seconds']) / max(rounds, 1)
    out = dict(config=a.config, s=a.s, n=c['n'], m=c['m'], N=c['N'], q=c['q'], p=c['p'], k=c['k'], B=c['B'],
               seconds=r['seconds'], bank_seconds=r['bank_seconds'], rounds=rounds, truncated=r['truncated'],
               per_round_seconds=per_round, spmv=r['spmv'], rhs=r['rhs'], sequence=r['sequence'],
               time_limit=a.time_limit, workers=a.workers, input_seconds=prep, cpu_model=cpu_model(),
               os_cpu_count=os.cpu_count(), affinity=len(os.sched_getaffinity(0)))
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
