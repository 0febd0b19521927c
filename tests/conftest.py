"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built library;
everything else runs on the CPU (oracle pins, loader, host logic, ABI)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def P():
    from contrafold_gpu_amd import params
    return params.load()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def engine(P):
    from contrafold_gpu_amd import default_engine
    eng = default_engine()
    eng.set_params(P)
    return eng


def rand_seq(rng, L, alphabet="ACGU", p=None):
    return "".join(rng.choice(list(alphabet), size=L, p=p))
