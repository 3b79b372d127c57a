import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP device (MI355X)')
    config.addinivalue_line('markers', 'slow: long-running test')


def load_golden(name):
    """Golden fixture written by tests/golden/make_golden.py (plain arrays, no pickle)."""
    z = np.load(os.path.join(GOLDEN, name + '.npz'), allow_pickle=False)
    out = {k: z[k] for k in z.files}

    def tt(prefix):
        return [out['%s_%d' % (prefix, i)] for i in range(int(out[prefix + '_order']))]

    out['tt'] = tt
    return out


@pytest.fixture
def golden():
    return load_golden
