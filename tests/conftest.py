"""Shared fixtures: golden vectors, the CPU oracle and (for ``gpu`` tests) the hip backend."""

from __future__ import annotations

import os
import sys
from types import SimpleNamespace

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libkrhip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.build()
    return O


def _load(name: str) -> dict:
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden1() -> dict:
    return _load("bardeen64_a0_i60.npz")


@pytest.fixture(scope="session")
def golden2() -> dict:
    return _load("bardeen512_a09_i60.npz")


@pytest.fixture(scope="session")
def golden_bisect() -> dict:
    return _load("bisect_a09_i60_16dirs.npz")


@pytest.fixture(scope="session")
def hip():
    """The hip backend; a GPU test fails loudly (never skips) if it cannot be opened."""
    from kerrray_amd.backends import HipBackend

    return HipBackend()


def options(**kw):
    """Reference-default integrator/termination options with overrides."""
    from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions

    iargs = {k: v for k, v in kw.items() if k in IntegratorOptions.__dataclass_fields__}
    targs = {k: v for k, v in kw.items() if k in TerminationOptions.__dataclass_fields__}
    return IntegratorOptions(**iargs), TerminationOptions(**targs)


def ns(**kw) -> SimpleNamespace:
    return SimpleNamespace(**kw)
