"""Shared fixtures: loaders for the product package (directory new-repo_amd/),
its host modules and the CPU oracle (test infrastructure)."""
import importlib.util
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "new-repo_amd"
ORACLE_DIR = ROOT / "oracle"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblesgpu.so)")


def _load(name, path, search=None):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path, submodule_search_locations=search)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_pkg():
    return _load("new_repo_amd", PKG_DIR / "__init__.py", [str(PKG_DIR)])


def load_params():
    return load_pkg().load_params_module()


def load_ic():
    return load_pkg().load_ic_module()


def load_oracle():
    return _load("les_oracle_py", ORACLE_DIR / "oracle.py")


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def params():
    return load_params()


@pytest.fixture(scope="session")
def icmod():
    return load_ic()


@pytest.fixture(scope="session")
def oracle():
    return load_oracle()


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num
