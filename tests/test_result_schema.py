"""run_real's time_*.json and effect_*.json carry every field of the reference's stored result files
(realworld/results/, key trees in tests/golden/result_schema.json; run_real.py:64-65, 117-118).
CPU test: the writer runs with the oracle standing in for the GPU greedy and evaluators, so the
check covers the file layout run_real builds; tests/test_gpu_run_real.py repeats it on the GPU path."""
import json

import numpy as np
import pytest

pytest.importorskip('torch')

from oracle import polar_oracle as O  # noqa: E402
from polar_amd import exact, run_real  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402
import result_schema as RS  # noqa: E402


def _oracle_graph(G):
    return O.Graph(G.n, G.edges, G.sigma, weights=None if G.unit else G.weights)


def _polar(G, cand, k, s, q=128, B=16, eps=1e-2, tol=1e-7, obj='frac', sketch=None, naive=False, time_limit=None,
           poles=None, **_):
    return O.polar(_oracle_graph(G), cand, k, s, q=q, B=B, eps=eps, tol=tol, obj=obj, sketch=sketch, naive=naive,
                   time_limit=time_limit, poles=poles)


def _multi(M0, cand, seq, targets):
    return [O.eval_prefix(M0, cand, seq, s, obj) for obj, s in targets]


def _true(G, cand, seq, s, obj='frac', **_):
    return O.true_marginals(_oracle_graph(G), cand, seq, s, obj=obj)


@pytest.fixture
def oracle_backend(monkeypatch):
    monkeypatch.setattr(pg, 'polar', _polar)
    monkeypatch.setattr(pg, 'true_marginals', _true)
    monkeypatch.setattr(exact, 'eval_prefix_multi', _multi)


def test_reference_schema_fixture_shapes():
    t = RS.reference('time')
    assert set(t) == {'mode', 'graph', 'params', 'env', 'runs'} and 'POLAR(B=16)' in t['runs']
    assert set(t['runs']['POLAR(B=16)'][0]) >= {'sequence', 'seconds', 'spmv', 'rhs', 'p'}
    e = RS.reference('effect_certified')
    assert set(e['gains']['frac(s=0.5)']['_evaluator']) == {'evaluator', 'quadrature_bound', 'residual_rho', 'p',
                                                             'spmv'}


def test_checker_reports_missing_and_mistyped_paths():
    ref = {'a': {'b': 'int', 'POLAR(B=16)': [{'c': 'str'}]}, 'env': {'cupy': 'str'}}
    assert RS.missing(ref, {'a': {'b': 'float', 'POLAR(B=3)': [{'c': 'str'}]}, 'env': {'torch': 'str'}}) == []
    assert RS.missing(ref, {'a': {'b': 'str', 'POLAR(B=3)': [{}]}, 'env': {}}) == [
        'a/b (int vs str)', 'a/POLAR(B=*)/[0]/c', 'env/cupy']


@pytest.mark.parametrize('dense_max', [6000, 10])
def test_time_and_effect_files_carry_the_reference_fields(tmp_path, oracle_backend, dense_max):
    common = ['--graphs', 'c1', '--N', '300', '--k', '3', '--q', '8', '--B', '2', '--s', '0.5', '--sigma', '0.1',
              '--poles', '8', '--out', str(tmp_path)]
    run_real.main(['time', *common, '--reps', '1'])
    d = json.loads((tmp_path / 'time_c1.json').read_text())
    assert RS.missing(RS.reference('time'), RS.tree(d)) == []
    run_real.main(['effect', *common, '--random', '2', '--dense-max', str(dense_max)])
    e = json.loads((tmp_path / 'effect_c1.json').read_text())
    kind = 'effect_dense' if dense_max > 2000 else 'effect_certified'
    # the reference's effect files also hold SLQ rows (run with --slq, a GPU baseline here) and, for
    # small graphs, its dense exact greedy (Exact*, realworld/exact.py; outside this path's scope)
    skip = ('SLQ(s=*)', 'Exact(s=*)', 'Exact-ER', 'Exact-Logdet')
    assert RS.missing(RS.reference(kind), RS.tree(e), skip=skip) == []
    ev = e['gains']['frac(s=0.5)']['_evaluator']['evaluator']
    assert ev == ('eigenvalues' if kind == 'effect_dense' else 'certified-quadrature')
    assert np.all(np.diff(e['gains']['frac(s=0.5)']['POLAR(s=0.5)']) > 0)
