"""GPU: the HIP path reproduces the committed fixtures of tests/golden (the
oracle's output, frozen) through the C ABI: IEEE FP64 within 1e-12 relative
L1 per conserved variable (bit for bit in practice) and the same time, fast
FP64 within 1e-12."""
import glob
import json
import os

import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, 'golden', '*.npz')))

NAMES = {'dim': 'dim', 'size': 'gridSize', 'mins': 'mins', 'maxs': 'maxs', 'plm': 'usePLM',
         'slope_limiter': 'slopeLimiter', 'flux_limiter': 'fluxLimiter', 'integrator': 'integrator',
         'cfl': 'cfl', 'boundary': 'boundary', 'boundary_fixed': 'boundaryFixed'}


@pytest.mark.parametrize('fast', [False, True])
@pytest.mark.parametrize('path', FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_fixture(path, fast):
    from hydro.solver.hipfvsolver import HipFVSolver
    with np.load(path, allow_pickle=False) as z:
        f = {k: z[k] for k in z.files}
    kw = json.loads(str(f['cfg']))
    s = HipFVSolver(fastMath=fast, **{NAMES[k]: v for k, v in kw.items()})
    s.resetState(f['U0'])
    s.update(int(f['steps']))
    dim = kw['dim']
    sl = [slice(None)] * 3
    for i in range(dim):
        sl[2 - i] = slice(2, -2)
    got = s.UBufObj.toCPU()[tuple(sl)][..., :5]
    ref = f['U'][tuple(sl)][..., :5]
    err = rel_l1(got, ref)
    print(os.path.basename(path), 'fast' if fast else 'ieee', err)
    assert np.all(err <= 1e-12), err
    if not fast:
        assert s.t == float(f['t'])
