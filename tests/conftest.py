"""Shared test setup.

Markers: ``gpu`` -- needs an AMD GPU and the built engine library; every other
test runs on the CPU (oracle, host logic, ABI exports).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
REFDATA = os.path.join(GOLDEN, "reference_data")
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built engine")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O  # oracle/oracle.py (test infrastructure)
    O.lib()
    return O


@pytest.fixture(scope="session")
def foxs():
    import pripps_amd
    return pripps_amd.Pripps(os.path.join(REFDATA, "foxs_formfactors.yaml"))


@pytest.fixture(scope="session")
def engine():
    from pripps_amd.engine import Engine
    eng = Engine(0)
    yield eng


def random_cluster(n, seed, radius=None, min_sep=1.4):
    """Small seeded globule (sequential random addition)."""
    from pripps_amd.synth import globule_positions
    rng = np.random.default_rng(seed)
    if radius is None:
        radius = max(3.0, 17.0 * (max(n, 1) / 1000.0) ** (1.0 / 3.0))
    return globule_positions(n, radius, rng) if n > 0 else np.zeros((0, 3))
