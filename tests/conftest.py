import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP) device')


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(autouse=True, scope='session')
def _log_home(tmp_path_factory):
    """Send the default '~/.mbirtorch' log paths into the pytest tmp tree."""
    base = tmp_path_factory.mktemp('mbir_home')
    real = os.path.expanduser
    prefix = os.path.join('~', '.mbirtorch')

    def expanduser(path):
        if isinstance(path, str) and (path == prefix or path.startswith(prefix + os.sep)):
            return os.path.join(str(base), path[2:])
        return real(path)

    with pytest.MonkeyPatch.context() as mp:
        mp.setattr(os.path, 'expanduser', expanduser)
        yield


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
