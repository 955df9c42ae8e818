import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / 'braids-hkt_amd'):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device) and libbraids_hip.so')


@pytest.fixture(scope='session', autouse=True)
def oracle_built():
    """The oracle library is tiny plain C; build it if a checkout lacks it."""
    lib = ROOT / 'oracle' / 'liboracle.so'
    if not lib.exists():
        subprocess.run(['make', '-s', '-C', str(ROOT / 'oracle')], check=True)
    return lib
