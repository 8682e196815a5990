"""Registers this directory as the package ``esys_particle_amd`` in a fresh
interpreter (worker processes started with the spawn method run this file
first, so that tasks pickled by reference to the package's functions load)."""
import importlib.util
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
if "esys_particle_amd" not in sys.modules:
    _spec = importlib.util.spec_from_file_location(
        "esys_particle_amd", os.path.join(_HERE, "__init__.py"), submodule_search_locations=[_HERE])
    _mod = importlib.util.module_from_spec(_spec)
    sys.modules["esys_particle_amd"] = _mod
    _spec.loader.exec_module(_mod)
