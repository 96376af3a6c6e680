import json
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def golden(name: str):
    return json.loads((GOLDEN / name).read_text())


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    O.set_threads(int(os.environ.get("WSR_ORACLE_THREADS", "4")))
    return O


@pytest.fixture(scope="session")
def mlp64():
    from weno_sr_publishing_amd.checkpoints import load_f64
    return load_f64()


@pytest.fixture(scope="session")
def mlp32():
    from weno_sr_publishing_amd.checkpoints import load_f32
    return load_f32()


@pytest.fixture(scope="session")
def gpu_device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def mlp7():
    from weno_sr_publishing_amd.checkpoints import load_weno7_checkpoint, WENO7_SR_FP64
    return load_weno7_checkpoint(WENO7_SR_FP64)
