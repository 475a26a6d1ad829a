import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs an AMD GPU (runs the hip backend)')


@pytest.fixture(scope='session')
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail('this test needs a visible GPU (run with -m "not gpu" '
                    'on a machine without one)')
    from easier_amd import _native
    _native.lib()
    return torch.device('cuda', 0)
