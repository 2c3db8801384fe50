import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def gpu_lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from damask_amd import _lib
    return _lib.load()


def fcc_golden_inputs():
    from damask_amd import config, geom
    grid = geom.read_vti(os.path.join(GOLDEN, "fcc_6grains6x7x8.vti"))
    mat = config.parse_material(open(os.path.join(GOLDEN, "fcc_singlePhase.yaml")).read())
    _, lcs = config.parse_load(open(os.path.join(GOLDEN, "fcc_tensionY.yaml")).read())
    return grid, mat, lcs


def bcc_golden_inputs():
    from damask_amd import config, geom
    grid = geom.read_vti(os.path.join(GOLDEN, "bcc_17grains.vti"))
    mat = config.parse_material(open(os.path.join(GOLDEN, "bcc_material.yaml")).read())
    _, lcs = config.parse_load(open(os.path.join(GOLDEN, "bcc_tensionX.yaml")).read())
    num = config.parse_numerics(open(os.path.join(GOLDEN, "bcc_numerics.yaml")).read())
    return grid, mat, lcs, num
