"""Shared pytest configuration.

`gpu` marks tests that need an MI355X (they call the HIP library through its C ABI); everything
else runs on the CPU. Helpers here locate the oracle and the product package by path so the
tests never depend on an installed copy.
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "lightweight-mb-color-gradient-lbm_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


def load_product():
    """Import the product package (directory name is not a Python identifier)."""
    if "cglbm" in sys.modules:
        return sys.modules["cglbm"]
    spec = importlib.util.spec_from_file_location("cglbm", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cglbm"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def oracle():
    import cgref
    if not os.path.exists(cgref.LIB_PATH):
        pytest.fail("oracle/liboracle.so missing: run __graft_entry__.build() or make -C oracle")
    return cgref
