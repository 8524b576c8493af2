"""The run_real drop-in CLI (time and effect modes) writes the reference's JSON schema."""
import json

import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def test_time_and_effect_modes(tmp_path):
    from polar_amd import run_real
    common = ['--graphs', 'c1', '--N', '800', '--k', '4', '--q', '16', '--B', '3', '--s', '0.5',
              '--sigma', '0.1', '--poles', '12', '--kind', 'rademacher', '--out', str(tmp_path)]
    run_real.main(['time', *common, '--reps', '1'])
    d = json.loads((tmp_path / 'time_c1.json').read_text())
    assert d['mode'] == 'time' and set(d['runs']) == {'Naive', 'POLAR(B=1)', 'POLAR(B=3)'}
    r = d['runs']
    assert r['Naive'][0]['sequence'] == r['POLAR(B=1)'][0]['sequence']
    assert r['POLAR(B=1)'][0]['rhs'] == 2 * 16 + 4 * 1 and r['Naive'][0]['rhs'] == 4 * 2 * 16
    for k in ('sequence', 'seconds', 'spmv', 'rhs', 'p', 'rounds', 'truncated'):
        assert k in r['POLAR(B=3)'][0]
    assert d['graph']['N'] == 800 and d['graph']['sigma'] == 0.1
    run_real.main(['effect', *common, '--random', '2', '--slq', '--slq-probes', '4', '--slq-steps', '10'])
    e = json.loads((tmp_path / 'effect_c1.json').read_text())
    assert {'POLAR(s=0.5)', 'Sketched-ER', 'Sketched-Logdet', 'Degree', 'Random#0', 'Random#1',
            'SLQ(s=0.5)'} <= set(e['selections'])
    assert len(e['gains']['frac(s=0.5)']['SLQ(s=0.5)']) == 4
    g = e['gains']['frac(s=0.5)']
    assert g['_evaluator'] == {'evaluator': 'eigenvalues'} and len(g['POLAR(s=0.5)']) == 4
    assert all(x > 0 for x in g['POLAR(s=0.5)'])
    assert g['POLAR(s=0.5)'][-1] >= g['Random#0'][-1]
    import result_schema as RS
    assert RS.missing(RS.reference('time'), RS.tree(d)) == []
    assert RS.missing(RS.reference('effect_dense'), RS.tree(e), skip=('Exact(s=*)', 'Exact-ER', 'Exact-Logdet')) == []


def test_time_mode_gradient_shortlist(tmp_path):
    """--shortlist gradient: the first-order gain ranks the shortlist; the re-rank by exact
    marginals is unchanged, so the run keeps the schema and the rhs counter."""
    from polar_amd import run_real
    run_real.main(['time', '--graphs', 'c1', '--N', '800', '--k', '4', '--q', '16', '--B', '3', '--s', '0.5',
                   '--sigma', '0.1', '--poles', '12', '--kind', 'rademacher', '--out', str(tmp_path),
                   '--reps', '1', '--only', 'B', '--shortlist', 'gradient'])
    d = json.loads((tmp_path / 'time_c1.json').read_text())
    assert d['params']['shortlist'] == 'gradient' and set(d['runs']) == {'POLAR(B=3)'}
    r = d['runs']['POLAR(B=3)'][0]
    assert r['rhs'] == 2 * 16 + 4 * 3 and len(r['sequence']) == 4 and len(set(r['sequence'])) == 4
