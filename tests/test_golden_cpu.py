"""The committed fixtures of tests/golden (made by make_golden.py from the
oracle after it passed its closed-form pins) stay reproducible: the initial
states, and the oracle's final states bit for bit."""
import glob
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, 'golden', '*.npz')))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize('path', FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_reproduces_fixture(path):
    import sys
    sys.path.insert(0, os.path.join(HERE, 'golden'))
    import make_golden
    f = load(path)
    name = os.path.basename(path)[:-4]
    case = [c for c in make_golden.cases() if c[0] == name][0]
    _, steps, kw, init = case
    assert json.loads(str(f['cfg'])) == json.loads(json.dumps(kw))
    U0, U, t = make_golden.solve(kw, init, steps)
    assert np.array_equal(U0, f['U0'])
    assert np.array_equal(U, f['U'])
    assert t == float(f['t'])


def test_fixtures_present():
    assert len(FIXTURES) == 3
