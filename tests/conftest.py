import json
import os
import sys
from pathlib import Path

# the CPU oracle runs its single-phase step loops on OpenMP threads; parallel
# test workers each keep to one
if 'PYTEST_XDIST_WORKER' in os.environ:
    os.environ.setdefault('OMP_NUM_THREADS', '1')

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / 'tests' / 'golden'


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device)')
    config.addinivalue_line('markers', 'slow: long CPU oracle run')


@pytest.fixture(scope='session')
def records():
    return json.loads((GOLDEN / 'reference_records.json').read_text())
