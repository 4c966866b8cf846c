"""Shared test set-up: the `gpu` marker, the package loader and the oracle path."""
import importlib.util
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "helm-project_amd"
sys.path.insert(0, str(ROOT / "oracle"))


def load_package():
    """Imports helm-project_amd/ as `helm_project_amd` (its directory name is not an identifier)."""
    if "helm_project_amd" in sys.modules:
        return sys.modules["helm_project_amd"]
    spec = importlib.util.spec_from_file_location("helm_project_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["helm_project_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
