"""Writes the committed golden fixtures under tests/golden/ from the reference's stored results.

* scalability_fixture.json -- synthetic/scalability/results.json (the reference's CPU fixture:
  Naive-ShiftedCG / Inc-Full / POLAR-Lazy sequences, spmv and norm-evaluation counts on seeded
  graphs).  First repetition of each method; the reference asserts all repetitions agree
  (synthetic/scalability/benchmark.py:140-141).
* realworld_meta.json -- from realworld/results/time_*.json and effect_com-Youtube.json: the
  graph meta (n, m, dmax, sigma), the run parameters (k, eps, s), the pole count `p` and `rhs`
  counters of each timed method, and the certified evaluator's `quadrature_bound`,
  `residual_rho`, `p`.  These pin the composite rule's pole count and smallest node and the
  `rhs` formulas without the (absent) SNAP datasets.

* result_schema.json -- the key trees (keys, nesting, value types; lists by their first element)
  of three of the reference's stored result files: time_com-Amazon.json (time mode),
  effect_ca-GrQc.json (effect mode, dense eigenvalue evaluator) and effect_com-Youtube.json
  (effect mode, certified quadrature evaluator).  run_real's output must contain every path.

Run where the reference is mounted:  python tests/golden/extract_fixtures.py /root/reference
"""
import glob
import json
import sys
from pathlib import Path


def main(ref_root):
    root = Path(ref_root)
    here = Path(__file__).resolve().parent
    d = json.loads((root / 'synthetic' / 'scalability' / 'results.json').read_text())
    out = dict(source='synthetic/scalability/results.json', parameters=d['parameters'], configurations=[])
    for c in d['configurations']:
        row = dict(n=c['n'], m=c['m'], topology=c['topology'], support=c['support'],
                   n_candidates_vertices=c['n_candidates_vertices'], seed=c['seed'], methods={})
        for name, runs in c['raw'].items():
            r = runs[0]
            row['methods'][name] = dict(sequence=r['sequence'], spmvs=r['spmvs'],
                                        norm_evaluations=r['norm_evaluations'])
        out['configurations'].append(row)
    (here / 'scalability_fixture.json').write_text(json.dumps(out, indent=1))

    meta = dict(source='realworld/results/time_*.json, effect_com-Youtube.json', time=[], evaluator=[])
    for f in sorted(glob.glob(str(root / 'realworld' / 'results' / 'time_*.json'))):
        t = json.loads(Path(f).read_text())
        g, pr = t['graph'], t['params']
        meta['time'].append(dict(graph=g['name'], n=g['n'], m=g['m'], dmax=g['dmax'], sigma=g['sigma'],
                                 k=pr['k'], eps=pr['eps'], s=pr['s'], q=pr['q'], B=pr['B'],
                                 p={k: v[0]['p'] for k, v in t['runs'].items()},
                                 rhs={k: v[0]['rhs'] for k, v in t['runs'].items()},
                                 rounds_naive=None))
    e = json.loads((root / 'realworld' / 'results' / 'effect_com-Youtube.json').read_text())
    for key, v in e['gains'].items():
        ev = v['_evaluator']
        meta['evaluator'].append(dict(graph='com-Youtube', dmax=e['graph']['dmax'], sigma=e['graph']['sigma'],
                                      k=e['params']['k'], target=key, quadrature_bound=ev['quadrature_bound'],
                                      residual_rho=ev['residual_rho'], p=ev['p']))
    (here / 'realworld_meta.json').write_text(json.dumps(meta, indent=1))

    def tree(x):
        if isinstance(x, dict):
            return {k: tree(v) for k, v in x.items()}
        if isinstance(x, list):
            return [tree(x[0])] if x else []
        return type(x).__name__
    schema = dict(source='realworld/results/{time_com-Amazon,effect_ca-GrQc,effect_com-Youtube}.json key trees')
    for key, f in (('time', 'time_com-Amazon.json'), ('effect_dense', 'effect_ca-GrQc.json'),
                   ('effect_certified', 'effect_com-Youtube.json')):
        schema[key] = tree(json.loads((root / 'realworld' / 'results' / f).read_text()))
    (here / 'result_schema.json').write_text(json.dumps(schema, indent=1))
    print('wrote fixtures to', here)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
