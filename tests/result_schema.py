"""Key-tree comparison of run_real's result files with the reference's stored ones
(tests/golden/result_schema.json, extracted from realworld/results/ by tests/golden/extract_fixtures.py).

A key tree maps dict keys to subtrees, a list to [its first element's tree] and a scalar to its type
name.  Our file must contain every path of the reference's tree.  Data-dependent labels are compared
by pattern (POLAR(B=16) ~ POLAR(B=3), frac(s=0.5) ~ frac(s=0.25), Random#7 ~ Random#0, ...), ints and
floats are both numbers, and two fields differ by design: env records `torch` where the reference
records `cupy` (there is no CuPy here), and graph.sha256 exists only for downloaded graphs."""
import json
import re
from pathlib import Path

GOLD = Path(__file__).resolve().parent / 'golden' / 'result_schema.json'
_PATTERNS = [(re.compile(r'^POLAR\(B=\d+\)$'), 'POLAR(B=*)'), (re.compile(r'^POLAR\(s=[\d.]+\)$'), 'POLAR(s=*)'),
             (re.compile(r'^SLQ\(s=[\d.]+\)$'), 'SLQ(s=*)'), (re.compile(r'^Exact\(s=[\d.]+\)$'), 'Exact(s=*)'), (re.compile(r'^frac\(s=[\d.]+\)$'), 'frac(s=*)'),
             (re.compile(r'^Random#\d+$'), 'Random#*')]
_NUM = {'int', 'float'}
_RENAMED = {('env', 'cupy'): 'torch'}


def reference(kind):
    return json.loads(GOLD.read_text())[kind]


def tree(x):
    if isinstance(x, dict):
        return {k: tree(v) for k, v in x.items()}
    if isinstance(x, list):
        return [tree(x[0])] if x else []
    return type(x).__name__


def _label(k):
    for rx, name in _PATTERNS:
        if rx.match(k):
            return name
    return k


def _norm(d):
    out = {}
    for k, v in d.items():
        out.setdefault(_label(k), v)
    return out


def missing(ref, ours, path=(), skip=()):
    """Paths of `ref` that `ours` lacks (or holds with another type), as 'a/b/c' strings."""
    if isinstance(ref, dict):
        if not isinstance(ours, dict):
            return ['/'.join(path) + ' (not a dict)']
        r, o = _norm(ref), _norm(ours)
        bad = []
        for k, v in r.items():
            if path[-1:] == ('graph',) and k == 'sha256' or k in skip:
                continue
            key = _RENAMED.get((path[-1] if path else '', k), k)
            if key not in o:
                bad.append('/'.join(path + (k,)))
            else:
                bad += missing(v, o[key], path + (k,), skip)
        return bad
    if isinstance(ref, list):
        if not isinstance(ours, list):
            return ['/'.join(path) + ' (not a list)']
        if ref and ours:
            return missing(ref[0], ours[0], path + ('[0]',), skip)
        return []
    if ref in _NUM and ours in _NUM or ref == ours or ref == 'NoneType' or ours == 'NoneType':
        return []
    return ['/'.join(path) + f' ({ref} vs {ours})']
