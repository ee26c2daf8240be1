import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / 'tests' / 'golden'

# The oracle's Krylov loops alternate its OpenMP kernel with short BLAS calls;
# a multithreaded BLAS pool spinning beside the OpenMP team on the same cores
# slows both several times over, so the BLAS side runs single-threaded.
import numpy  # noqa: E402,F401  (load both BLAS builds before limiting them)
import scipy.linalg  # noqa: E402,F401
from threadpoolctl import threadpool_limits  # noqa: E402

_BLAS_LIMIT = threadpool_limits(1, user_api='blas')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP device (MI355X) and the built engine library')


@pytest.fixture(scope='session')
def golden_dir():
    return GOLDEN


@pytest.fixture(scope='session')
def device():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('gpu-marked test run without a HIP device')
    return torch.device('cuda:0')
