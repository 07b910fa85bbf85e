import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, 'hydro-cl-lua_amd'), os.path.join(REPO, 'oracle'), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs through the HIP C ABI)')


def rel_l1(got, ref, axis=None):
    """relative L1 error per conserved variable (last axis)"""
    import numpy as np
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    ax = tuple(range(ref.ndim - 1)) if axis is None else axis
    den = np.abs(ref).sum(axis=ax)
    num = np.abs(got - ref).sum(axis=ax)
    return num / np.where(den == 0, 1, den)
