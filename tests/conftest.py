import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs an MI355X (HIP device); run with -m gpu')


@pytest.fixture(scope='session')
def gpu_ctx_factory():
    from pyxopto_amd import _native as N
    if N.device_count() < 1:
        pytest.fail('no HIP device visible for a gpu test')
    return N.Context
