"""Shared test setup: the ``gpu`` marker, the package loader and the oracle.

The package directory (geos-chem-gpu-hybrid_amd) is not a Python identifier;
it is loaded under the name ``gcchem_amd``.  The CPU oracle (oracle/) is the
checker of every parity test.
"""
import gzip
import importlib.util
import json
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
PKG = REPO / "geos-chem-gpu-hybrid_amd"


def load_package():
    if "gcchem_amd" in sys.modules:
        return sys.modules["gcchem_amd"]
    spec = importlib.util.spec_from_file_location(
        "gcchem_amd", PKG / "__init__.py", submodule_search_locations=[str(PKG)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gcchem_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()
sys.path.insert(0, str(REPO / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def mech():
    from gcchem_amd import mechanism
    return mechanism.load("fullchem")


@pytest.fixture(scope="session")
def oracle_libm(mech):
    from oracle import Oracle
    return Oracle(mech.ir, pow_mode="libm")


@pytest.fixture(scope="session")
def oracle_cr(mech):
    from oracle import Oracle
    return Oracle(mech.ir, pow_mode="cr")


@pytest.fixture(scope="session")
def checkforms():
    with gzip.open(REPO / "tests" / "golden" / "fullchem_checkforms.json.gz", "rt") as f:
        return json.load(f)
