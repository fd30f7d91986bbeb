"""Shared fixtures. Tests marked `gpu` need an AMD GPU and the built libx3dpt.so; the rest run on
the CPU (oracle against the reference build and golden fixtures, host logic, ABI exports)."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running parity test")


@pytest.fixture(scope="session")
def oracle_mod():
    import bindings

    if not bindings.oracle_available():
        pytest.skip("oracle/liboracle.so not built (run `make -C oracle`)")
    return bindings


@pytest.fixture(scope="session")
def ref_mod():
    import bindings

    if not bindings.reference_available():
        pytest.skip("oracle/_ref/libx3dref.so not built (needs the reference sources: `make -C oracle ref`)")
    return bindings


@pytest.fixture(scope="session")
def x3d():
    import x3dcore2_amd

    x3dcore2_amd.lib()  # raises loudly if the HIP library is missing
    return x3dcore2_amd
