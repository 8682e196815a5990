"""The C-ABI library loads and exports every entry point include/esp.h declares
(no compute call: runs without a GPU), and the oracle exports the same entry
points with its eso_ prefix."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "esp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(esp_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("esp_create", "esp_destroy", "esp_add_particles", "esp_add_rot_bond_group",
                 "esp_add_rot_friction_group", "esp_add_nrot_elastic_group",
                 "esp_run_time_steps", "esp_get_particle_data", "esp_last_error"):
        assert must in names


def test_hip_library_exports_every_symbol(pkg):
    if not os.path.exists(pkg.LIB_PATH):
        pytest.skip("libesp.so not built")
    dll = ctypes.CDLL(pkg.LIB_PATH)
    missing = [n for n in declared() if not hasattr(dll, n)]
    assert not missing, missing


def test_oracle_exports_the_same_entry_points(pkg, olib):
    hip_only = {"esp_" + n for n in pkg._native.HIP_ONLY_SIGNATURES}
    missing = [n for n in declared()
               if n not in hip_only and not hasattr(olib.dll, "eso_" + n[len("esp_"):])]
    assert not missing, missing


def test_native_binding_covers_the_header(pkg):
    bound = {"esp_" + n for n in list(pkg._native.SIGNATURES) + list(pkg._native.HIP_ONLY_SIGNATURES)}
    assert set(declared()) == bound
