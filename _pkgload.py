"""Registers the package directory ``compute-vp9_amd`` as ``compute_vp9_amd``.

The directory name carries a hyphen, so it cannot be imported by name; this
helper loads its ``__init__.py`` under an importable module name.
"""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "compute-vp9_amd")


def load_package():
    if "compute_vp9_amd" in sys.modules:
        return sys.modules["compute_vp9_amd"]
    spec = importlib.util.spec_from_file_location(
        "compute_vp9_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["compute_vp9_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
