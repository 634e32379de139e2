import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP kernels through the C ABI)')


@pytest.fixture(scope='session')
def golden():
    import json
    g = os.path.join(HERE, 'golden')

    class G:
        dir = g

        @staticmethod
        def load(name):
            with open(os.path.join(g, name)) as f:
                return json.load(f)

        @staticmethod
        def circuit(name):
            with open(os.path.join(g, 'circuits', name + '.stim')) as f:
                return f.read()
    return G
