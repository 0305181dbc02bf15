"""Loads the product package from its directory (the name is not a Python
identifier)."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "lbgpu-prime-dev_amd")


def load():
    name = "lbgpu_prime_dev_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_submodule(name):
    """a module of the package (e.g. "deck")"""
    import importlib
    load()
    return importlib.import_module("lbgpu_prime_dev_amd." + name)
